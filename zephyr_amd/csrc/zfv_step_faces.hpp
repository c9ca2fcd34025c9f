// Per-face terms of the MUSCL step shared by the step kernels of
// zfv_kernels.hip (one thread per cell over the mesh arrays) and the tiled
// kernel of zfv_tile.hip (the same terms from values staged in LDS): the
// LSM and limiting terms of the gradient, the predictor's face flux and the
// corrector's face flux, each a function of the values it is given, so both
// paths make the same operations in the same order.
#pragma once

#include "zfv_internal.hpp"

namespace zfv {

// drn = (dr.n) n of the LSM (gradient.h:254); a canonical normal keeps one
// component of dr exactly.
__device__ __forceinline__ void lsm_drn(const Frame& f, double drx, double dry, double& dnx, double& dny) {
    if (!f.gen) {
        dnx = f.side < 2 ? drx : 0.0;
        dny = f.side < 2 ? 0.0 : dry;
        return;
    }
    double d = drx * f.nx + dry * f.ny;
    dnx = d * f.nx;
    dny = d * f.ny;
}

// FaceExtraDirect::m/p (math/cfd/face_extra.h:143-177) on {rho, u, v, p};
// the energy is then recomputed from the EOS by every caller.
__device__ __forceinline__ Z extrap(const Z& z, const double* gx, const double* gy, double dx, double dy) {
    Z r;
    r.r = z.r + (gx[0] * dx + gy[0] * dy);
    r.u = z.u + (gx[1] * dx + gy[1] * dy);
    r.v = z.v + (gx[2] * dx + gy[2] * dy);
    r.p = z.p + (gx[3] * dx + gy[3] * dy);
    r.e = 0.0;
    return r;
}

// LSM data of one slot for the one-thread-per-cell gradient kernel:
// differences dF of {rho, u, v, p}, offset dr, its normal part drn, weight,
// and which rows of the LSM system the slot feeds (a canonical L/R normal
// has drn.y = 0, B/T drn.x = 0).
struct SlotL {
    double dF[4];
    double drx, dry, dnx, dny, w;
    bool xp, yp;
};

// LSM pass-1 terms: A += w drn dr^T, F += w dU drn (gradient.h:251-266).
__device__ __forceinline__ void lsm_add(const SlotL& L, double& A00, double& A01, double& A10, double& A11,
                                        double* Fx, double* Fy) {
    if (L.xp) {
        const double wd = L.w * L.dnx;
        A00 = A00 + wd * L.drx;
        A01 = A01 + wd * L.dry;
#pragma unroll
        for (int k = 0; k < 4; ++k) Fx[k] = Fx[k] + (L.w * L.dF[k]) * L.dnx;
    }
    if (L.yp) {
        const double wd = L.w * L.dny;
        A10 = A10 + wd * L.drx;
        A11 = A11 + wd * L.dry;
#pragma unroll
        for (int k = 0; k < 4; ++k) Fy[k] = Fy[k] + (L.w * L.dF[k]) * L.dny;
    }
}

// Limiting terms (gradient.h:412-474): limiter(2 du_n - dF, dF) dF.
template <int LIM>
__device__ __forceinline__ void lim_add(const SlotL& L, const double* gx, const double* gy, double* Lx, double* Ly) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double du_n = gx[k] * L.drx + gy[k] * L.dry;
        const double dFm = 2.0 * du_n - L.dF[k];
        const double lim = limiter2<LIM>(dFm, L.dF[k]) * L.dF[k];
        if (L.xp) Lx[k] = Lx[k] + (L.w * lim) * L.dnx;
        if (L.yp) Ly[k] = Ly[k] + (L.w * lim) * L.dny;
    }
}

// Predictor flux term of a slot, flux * area (sm_fluid.cpp:176-253), from
// the face centre offset (dx, dy), the area and the frame.
__device__ __forceinline__ Q pred_eval(const Frame& fr, double dx, double dy, double a, const Z& zc,
                                       const double* gx, const double* gy, const Eos& eos) {
    Z zm = extrap(zc, gx, gy, dx, dy);
    zm.e = eos_e(eos, zm.r, zm.p);
    if (is_bad(zm, eos)) zm = zc;
    Q f = phys_flux(frame_local(fr, zm));
    f = frame_global(fr, f);
    return Q{f.r * a, f.mx * a, f.my * a, f.E * a};
}

// Everything one corrector face reads: the slot's geometry and frame, and
// the neighbour's half state, gradient, centre and isolation counter.  On
// boundary slots adj is the cell itself, so the neighbour loads stay in
// bounds and can be issued before any choice is made.
struct FaceC {
    Z hn;
    double ngx[4], ngy[4];
    double ncx, ncy, fx, fy, a;
    Frame fr;
    int wnb;
};

// One face of SmFluid::fluxes_stage2 (sm_fluid.cpp:255-366): flux * area of
// a slot from its loads.  The isolation, boundary and bad-state choices are
// selects among computed states.
template <int FLUX>
__device__ __forceinline__ Q corr_eval(const FaceC& F, int nib, const Z& zch, int wc, double ccx, double ccy,
                                       const double* gx, const double* gy, const Eos& eos, bool& bad) {
    const int t = nib & 7;
    const Z& hn = F.hn;
    const int wnb = F.wnb;
    const double ncx = F.ncx, ncy = F.ncy, fx = F.fx, fy = F.fy, a = F.a;
    const double* ngx = F.ngx;
    const double* ngy = F.ngy;
    const Frame& fr = F.fr;
    const bool bound = t >= SC_WALL;
    const bool wall = t == SC_WALL;
    const bool robust = wc > 0 || wnb > 0;

    // face_extra: own side from the cell's half state, neighbour side from
    // the neighbour's (or, on a periodic slot, from the mirror point)
    Z zm = extrap(zch, gx, gy, fx - ccx, fy - ccy);
    zm.e = eos_e(eos, zm.r, zm.p);
    if (is_bad(zm, eos)) zm = zch;
    double px = ncx, py = ncy;
    if (t != SC_INNER) frame_symm(fr, ccx, ccy, fx, fy, px, py);
    Z zp = extrap(hn, ngx, ngy, fx - px, fy - py);
    zp.e = eos_e(eos, zp.r, zp.p);
    if (is_bad(zp, eos)) zp = hn;
    if (bound) zp = frame_boundary(fr, zm, wall);
    if (robust) {
        zm = zch;
        zp = bound ? frame_boundary(fr, zch, wall) : hn;
    }
    const Z lm = frame_local(fr, zm), lp = frame_local(fr, zp);
    Q f = robust ? flux_hll(lm, lp, eos, bad) : num_flux<FLUX>(lm, lp, eos, bad);
    f = frame_global(fr, f);
    return Q{f.r * a, f.mx * a, f.my * a, f.E * a};
}

}  // namespace zfv
