// Transfer (math/solver/transfer.h:21-153, transfer.cpp:1-719): a volume
// fraction u carried by a given velocity field, one thread per cell on the
// context's mesh (mesh order).  The velocity is the caller's samples at the
// face centres (fvx, fvy, slot-major [8][stride]) and cell centres (cvx,
// cvy), the values the reference's virtual velocity() returns there.
//
// Kernels: the slopes of the MUSCL methods (gradient::LSM<double> and, for
// MUSCL_MC, gradient::limiting with MC, boundary value = u), the face fluxes
// and u_new = between(u - sum F / V, 0, 1) into the other layer, the
// interface normals (InterfaceRecovery::compute_normal), the CFL reduction
// and set_flags.  Built methods, Transfer::Method numbering: CRP_V3 0,
// CRP_V5 1, CRP_SE 2, CRP_N1 3, MUSCLd 7, MUSCLd_CRP 9, MUSCL_MC 11,
// MUSCL_MC_CRP 12, WENO 13 and WENO_CRP 14 (on a Rectangle's grid).
#include "zfv_internal.hpp"
#include "zfv_step_common.hpp"
#include "zfv_step_faces.hpp"

namespace zfv {

using namespace mm;

namespace {

constexpr int kTrV3 = 0, kTrV5 = 1, kTrSE = 2, kTrN1 = 3, kTrMUSCLd = 7, kTrMUSCLdCRP = 9, kTrMC = 11, kTrMCCRP = 12,
              kTrWENO = 13, kTrWENOCRP = 14;
constexpr int kLimMC = 3, kLimVanAlbada2 = 6;   // limiter ids (the MC of m_limiter, ATvL's vanAlbada2)

__device__ __forceinline__ bool tr_swept(const DevMesh& M, int so, int dir) {   // EuCell::faces(dir)
    if (dir == 1) return !(fabs(M.nx[so]) < 0.7);
    if (dir == 2) return !(fabs(M.ny[so]) < 0.7);
    return true;
}

__device__ __forceinline__ double tr_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// flux_2D (transfer.cpp:110-126).
__device__ __forceinline__ double tr_flux_2D(double a1, double a2, double S, double V1, double V2, double as,
                                             double vn, double dt) {
    const double a = vn > 0.0 ? a1 : a2;
    const double V = vn > 0.0 ? V1 : V2;
    const double gamma = dt * fabs(vn) * S;
    const double F_min = smax(0.0, gamma - (1.0 - a) * V);
    const double F_max = a * V;
    return tr_sign(vn) * f_between(gamma * as, F_min, F_max);
}

// best_face_fraction + flux_CRP (transfer.cpp:183-206).
__device__ __forceinline__ double tr_flux_CRP(double a1, double a2, double S, double V1, double V2, double vn,
                                              double dt, double Flux) {
    double a_sig;
    if (vn == 0.0) {
        a_sig = 0.5 * (a1 + a2);
    } else {
        const double lo = a1 < a2 ? a1 : a2, hi = a1 < a2 ? a2 : a1;   // sorted
        a_sig = f_between(Flux / (dt * vn * S), lo, hi);
    }
    return tr_flux_2D(a1, a2, S, V1, V2, a_sig, vn, dt);
}

// One slot of the scalar LSM: the neighbour value (the cell's own at a
// boundary), dr to the neighbour centre or the mirror point, drn, weight.
struct TrSlot {
    double dF, drx, dry, dnx, dny, w;
    bool xp, yp;
};

__device__ __forceinline__ TrSlot tr_slot(const DevMesh& M, const double* __restrict__ u, int c, int s, int nib,
                                          double uc, double ccx, double ccy) {
    TrSlot L;
    const int so = s * M.stride + c;
    const int t = nib & 7;
    const Frame fr = slot_frame(M, c, s, nib);
    const int nb = M.adj[so];
    const double un = (t == SC_INNER || t == SC_PERIODIC) ? u[nb] : uc;
    double px = M.cx[nb], py = M.cy[nb];
    if (t != SC_INNER) frame_symm(fr, ccx, ccy, M.fcx[so], M.fcy[so], px, py);
    L.drx = px - ccx;
    L.dry = py - ccy;
    lsm_drn(fr, L.drx, L.dry, L.dnx, L.dny);
    L.w = M.w[so];
    L.dF = un - uc;
    L.xp = fr.gen || fr.side < 2;
    L.yp = fr.gen || fr.side >= 2;
    return L;
}

// Transfer::compute_slopes for MUSCLd / MUSCL_MC (transfer.cpp:453-472):
// the LSM gradient of u (the 2x2 block of the reference's 3x3 system, as in
// the MUSCL step), limited with LIM when LIMIT.
template <int LIM, bool LIMIT>
__global__ void __launch_bounds__(256) k_tr_slopes(DevMesh M, const double* __restrict__ u, double* __restrict__ gx_out,
                                                   double* __restrict__ gy_out) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= M.n) return;
    const double uc = u[c];
    const double ccx = M.cx[c], ccy = M.cy[c];
    const int code = M.code[c];
    double A00 = 0.0, A01 = 0.0, A10 = 0.0, A11 = 0.0, Fx = 0.0, Fy = 0.0;
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {
        const int nib = (code >> (4 * s)) & 15;
        if ((nib & 7) == SC_UNDEF) continue;
        const TrSlot L = tr_slot(M, u, c, s, nib, uc, ccx, ccy);
        if (L.xp) {
            const double wd = L.w * L.dnx;
            A00 = A00 + wd * L.drx;
            A01 = A01 + wd * L.dry;
            Fx = Fx + (L.w * L.dF) * L.dnx;
        }
        if (L.yp) {
            const double wd = L.w * L.dny;
            A10 = A10 + wd * L.drx;
            A11 = A11 + wd * L.dry;
            Fy = Fy + (L.w * L.dF) * L.dny;
        }
    }
    if (fabs(A11) < 1e-8) A11 = 1.0;
    const double det = A11 * A00 + (-A01) * A10;
    const double invdet = 1.0 / det;
    const double B00 = A11 * invdet, B01 = (-A01) * invdet;
    const double B10 = (-A10) * invdet, B11 = A00 * invdet;
    double gx = B00 * Fx + B01 * Fy;
    double gy = B10 * Fx + B11 * Fy;
    if (LIMIT) {   // gradient::limiting (gradient.h:412-474)
        double Lx = 0.0, Ly = 0.0;
#pragma unroll 1
        for (int s = 0; s < 8; ++s) {
            const int nib = (code >> (4 * s)) & 15;
            if ((nib & 7) == SC_UNDEF) continue;
            const TrSlot L = tr_slot(M, u, c, s, nib, uc, ccx, ccy);
            const double du_n = gx * L.drx + gy * L.dry;
            const double dFm = 2.0 * du_n - L.dF;
            const double lim = limiter2<LIM>(dFm, L.dF) * L.dF;
            if (L.xp) Lx = Lx + (L.w * lim) * L.dnx;
            if (L.yp) Ly = Ly + (L.w * lim) * L.dny;
        }
        gx = B00 * Lx + B01 * Ly;
        gy = B10 * Lx + B11 * Ly;
    }
    gx_out[c] = gx;
    gy_out[c] = gy;
}

// fluxes_CRP (transfer.cpp:208-251) / fluxes_MUSCL (:331-395), then the
// layer update of update_* (:493-535): out = between(u - sum F / V, 0, 1).
template <int METHOD>
__global__ void __launch_bounds__(256) k_tr_flux(DevMesh M, const double* __restrict__ u, const double* __restrict__ tnx,
                                                 const double* __restrict__ tny, const double* __restrict__ gx,
                                                 const double* __restrict__ gy, const double* __restrict__ fvx,
                                                 const double* __restrict__ fvy, double dt, int dir,
                                                 double* __restrict__ out) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= M.n) return;
    const int st = M.stride;
    const double a1 = u[c];
    const double V1 = M.vol[c];
    const double ccx = M.cx[c], ccy = M.cy[c];
    const int code = M.code[c];
    double fluxes = 0.0;
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {
        const int t = (code >> (4 * s)) & 7;
        const int so = s * st + c;
        if (t == SC_UNDEF || t > SC_PERIODIC || !tr_swept(M, so, dir)) continue;
        const int nb = M.adj[so];
        const double a2 = u[nb];
        const double fnx = M.nx[so], fny = M.ny[so];
        const double vn = (fvx[so] * fnx + fvy[so] * fny) + 0.0;
        const double S = M.area[so];
        const double V2 = M.vol[nb];
        if (METHOD <= kTrN1) {
            double a_sig;
            if (METHOD == kTrV3) {
                a_sig = face_fraction_v3(a1, a2);
            } else if (METHOD == kTrV5) {
                a_sig = face_fraction_v5(a1, a2);
            } else if (METHOD == kTrN1) {   // face_fraction_n1 (transfer.cpp:131-143)
                const double a = vn > 0.0 ? a1 : a2;
                const double V = vn > 0.0 ? V1 : V2;
                const double cs = vn > 0.0 ? (tnx[c] * fnx + tny[c] * fny) + 0.0 : -((tnx[nb] * fnx + tny[nb] * fny) + 0.0);
                const double C = dt * fabs(vn) * S / V;
                a_sig = f_between(average_flux(a, cs, C), a1, a2);
            } else {
                a_sig = face_fraction_s(a1, a2);
            }
            fluxes += tr_flux_2D(a1, a2, S, V1, V2, a_sig, vn, dt);
            continue;
        }
        const double fcx = M.fcx[so], fcy = M.fcy[so];
        const double ncx = M.cx[nb], ncy = M.cy[nb];
        double a_sig;
        if (METHOD == kTrMC || METHOD == kTrMCCRP) {   // FaceExtra::Direct (face_extra.h:135-177)
            const double du_c = (gx[c] * (fcx - ccx) + gy[c] * (fcy - ccy)) + 0.0 * 0.0;
            const double du_n = (gx[nb] * (fcx - ncx) + gy[nb] * (fcy - ncy)) + 0.0 * 0.0;
            a_sig = vn > 0.0 ? a1 + du_c : a2 + du_n;
            a_sig = f_between(a_sig, a1, a2);
        } else {   // FaceExtra::ATvL (face_extra.h:264-339)
            const double rcx = fcx - ccx, rcy = fcy - ccy;
            const double rnx = ncx - fcx, rny = ncy - fcy;
            const double dx_c = sqrt((rcx * rcx + rcy * rcy) + 0.0 * 0.0);
            const double dx_n = sqrt((rnx * rnx + rny * rny) + 0.0 * 0.0);
            const double du_c = ((gx[c] * rcx + gy[c] * rcy) + 0.0 * 0.0) / dx_c;
            const double du_n = ((gx[nb] * rnx + gy[nb] * rny) + 0.0 * 0.0) / dx_n;
            const double dx = dx_c + dx_n;
            if (vn > 0.0) {
                const double delta = dx_c / dx;
                const double eta = (12.0 * delta * delta - 1.0) / 24.0;
                const double dzp = a2 - a1;
                const double dzm = 2.0 * du_c * dx - dzp;
                const double sl = limiter2<kLimVanAlbada2>(dzm, dzp);
                a_sig = a1 + sl * ((0.5 * delta - sl * eta) * dzm + (0.5 * delta + sl * eta) * dzp);
            } else {
                const double delta = dx_n / dx;
                const double eta = (12.0 * delta * delta - 1.0) / 24.0;
                const double dzm = a2 - a1;
                const double dzp = 2.0 * du_n * dx - dzm;
                const double sl = limiter2<kLimVanAlbada2>(dzm, dzp);
                a_sig = a2 - sl * ((0.5 * delta + sl * eta) * dzm + (0.5 * delta - sl * eta) * dzp);
            }
        }
        double Flux = a_sig * vn * dt * S;
        if (METHOD == kTrMCCRP || METHOD == kTrMUSCLdCRP) Flux = tr_flux_CRP(a1, a2, S, V1, V2, vn, dt, Flux);
        fluxes += Flux;
    }
    out[c] = f_between(a1 - fluxes / V1, 0.0, 1.0);
}

// WENO<b, e> (math/calc/weno.h:39-120): the quadratic interpolant's face
// values of three cell averages, p = u(+1/2) and m = u(-1/2), as
// w1 a + w2 b + w3 c left to right, and the smoothness indicator
// 13/12 d2^2 + 1/4 d1^2.  S is the sub-stencil's first offset (-2, -1, 0).
template <int S, bool RIGHT>
__device__ __forceinline__ double weno3(double a, double b, double c) {
    if (RIGHT) {
        if (S == -2) return ((1.0 / 3.0) * a + (-7.0 / 6.0) * b) + (11.0 / 6.0) * c;
        if (S == -1) return ((-1.0 / 6.0) * a + (5.0 / 6.0) * b) + (1.0 / 3.0) * c;
        return ((1.0 / 3.0) * a + (5.0 / 6.0) * b) + (-1.0 / 6.0) * c;
    }
    if (S == -2) return ((-1.0 / 6.0) * a + (5.0 / 6.0) * b) + (1.0 / 3.0) * c;
    if (S == -1) return ((1.0 / 3.0) * a + (5.0 / 6.0) * b) + (-1.0 / 6.0) * c;
    return ((11.0 / 6.0) * a + (-7.0 / 6.0) * b) + (1.0 / 3.0) * c;
}
template <int S>
__device__ __forceinline__ double weno_beta(double a, double b, double c) {
    const double d1 = S == -2 ? (a - 4.0 * b) + 3.0 * c : (S == -1 ? a - c : (3.0 * a - 4.0 * b) + c);
    const double d2 = (a - 2.0 * b) + c;
    return (13.0 / 12.0) * d2 * d2 + (1.0 / 4.0) * d1 * d1;
}

// WENO5::p() / m() (weno.h:126-170) on u[0..4]: sub-stencils {u0 u1 u2}
// (offset -2), {u1 u2 u3} (-1), {u2 u3 u4} (0), linear weights 1/10, 6/10,
// 3/10 in that order for p (3/10, 6/10, 1/10 for m) over (1e-8 + beta)^2,
// normalised.
template <bool RIGHT>
__device__ __forceinline__ double weno5(const double* u) {
    const double b0 = weno_beta<0>(u[2], u[3], u[4]);
    const double b1 = weno_beta<-1>(u[1], u[2], u[3]);
    const double b2 = weno_beta<-2>(u[0], u[1], u[2]);
    const double d0c = RIGHT ? 3.0 / 10.0 : 1.0 / 10.0, d1c = 6.0 / 10.0, d2c = RIGHT ? 1.0 / 10.0 : 3.0 / 10.0;
    const double e0 = 1.0e-8 + b0, e1 = 1.0e-8 + b1, e2 = 1.0e-8 + b2;
    const double d0 = d0c / (e0 * e0), d1 = d1c / (e1 * e1), d2 = d2c / (e2 * e2);
    const double w0 = d0 / (d0 + d1 + d2), w1 = d1 / (d0 + d1 + d2), w2 = d2 / (d0 + d1 + d2);
    return (w0 * weno3<0, RIGHT>(u[2], u[3], u[4]) + w1 * weno3<-1, RIGHT>(u[1], u[2], u[3])) +
           w2 * weno3<-2, RIGHT>(u[0], u[1], u[2]);
}

// update_WENO (transfer.cpp:541-658) on a Rectangle's grid (cell = ny i + j,
// indices wrapped periodically as EuMesh::operator()(i, j) does): for each
// swept side L, R, B, T -- boundary faces included, with the wrapped
// stencil -- the WENO5 face value of the upwind cell (the cell itself when
// vn > 0, else the neighbour across the side) on its five cells along the
// side's axis, m() on its lower face and p() on its upper one, bounded to
// [0, 1]; flux a_sig vn dt S, corrected by flux_CRP for WENO_CRP; then
// out = between(u - sum F / V, 0, 1).
template <bool CRP>
__global__ void __launch_bounds__(256) k_tr_weno(DevMesh M, const double* __restrict__ u, const double* __restrict__ fvx,
                                                 const double* __restrict__ fvy, double dt, int dir, int gnx, int gny,
                                                 double* __restrict__ out) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= M.n) return;
    const int st = M.stride;
    const int i = c / gny, j = c - (c / gny) * gny;
    const double a1 = u[c];
    const double V1 = M.vol[c];
    double fluxes = 0.0;
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        if ((s < 2 && dir == 2) || (s >= 2 && dir == 1)) continue;
        const int so = s * st + c;
        const double vn = (fvx[so] * M.nx[so] + fvy[so] * M.ny[so]) + 0.0;
        const bool upper = (s & 1) != 0;
        const int k = s < 2 ? i : j;
        const int K = vn > 0.0 ? k : (upper ? k + 1 : k - 1);
        double sten[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int kk = K - 2 + q;
            const int ii = s < 2 ? (kk + gnx) % gnx : i, jj = s < 2 ? j : (kk + gny) % gny;
            sten[q] = u[gny * ii + jj];
        }
        const bool right = upper ? vn > 0.0 : !(vn > 0.0);
        const double a_sig = f_between(right ? weno5<true>(sten) : weno5<false>(sten), 0.0, 1.0);
        const double S = M.area[so];
        double Flux = a_sig * vn * dt * S;
        if (CRP) {
            const int nb = M.adj[so];
            Flux = tr_flux_CRP(a1, u[nb], S, V1, M.vol[nb], vn, dt, Flux);
        }
        fluxes += Flux;
    }
    out[c] = f_between(a1 - fluxes / V1, 0.0, 1.0);
}

// InterfaceRecovery::compute_normal (interface_recovery.cpp:17-39): zero in
// a pure cell, else -sum face_fraction(uc, un) S n_f over the inner faces,
// normalised.
__global__ void __launch_bounds__(256) k_tr_normals(DevMesh M, const double* __restrict__ u, double* __restrict__ onx,
                                                    double* __restrict__ ony) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= M.n) return;
    const double uc = u[c];
    double nx = 0.0, ny = 0.0;
    if (!(uc < 1.0e-12 || uc > 1.0 - 1.0e-12)) {
        const int code = M.code[c];
#pragma unroll 1
        for (int s = 0; s < 8; ++s) {
            const int t = (code >> (4 * s)) & 7;
            if (t == SC_UNDEF || t > SC_PERIODIC) continue;
            const int so = s * M.stride + c;
            const double un = u[M.adj[so]];
            const double Sx = M.nx[so] * M.area[so], Sy = M.ny[so] * M.area[so];
            const double ff = face_fraction(uc, un);
            nx = nx - ff * Sx;
            ny = ny - ff * Sy;
        }
        const double zz = (nx * nx + ny * ny) + 0.0 * 0.0;
        if (zz > 0.0) {
            const double r = sqrt(zz);
            nx = nx / r;
            ny = ny / r;
        }
    }
    onx[c] = nx;
    ony[c] = ny;
}

// Transfer::compute_dt per cell (transfer.cpp:86-93), V / max face area /
// |v(centre)|; the minimum over the launch into *slot.
__global__ void __launch_bounds__(256) k_tr_dt(DevMesh M, const double* __restrict__ cvx, const double* __restrict__ cvy,
                                               unsigned long long* slot) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    double cand = INFINITY;
    if (c < M.n) {
        const int code = M.code[c];
        double max_area = 0.0;
#pragma unroll 1
        for (int s = 0; s < 8; ++s)
            if (((code >> (4 * s)) & 7) != SC_UNDEF) max_area = smax(max_area, M.area[s * M.stride + c]);
        const double dx = M.vol[c] / max_area;
        const double vx = cvx[c], vy = cvy[c];
        cand = dx / sqrt((vx * vx + vy * vy) + 0.0 * 0.0);
    }
    block_min_atomic(slot, cand);
}

// Transfer::set_flags (transfer.cpp:668-690).
__global__ void __launch_bounds__(256) k_tr_flags(DevMesh M, const double* __restrict__ u, int* __restrict__ flag) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= M.n) return;
    double lo = u[c], hi = u[c];
    const int code = M.code[c];
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {
        const int t = (code >> (4 * s)) & 7;
        if (t == SC_UNDEF || t > SC_PERIODIC) continue;
        const double un = u[M.adj[s * M.stride + c]];
        lo = smin(lo, un);
        hi = smax(hi, un);
    }
    flag[c] = hi - lo > 0.0 ? 1 : -1;
}

inline dim3 tr_grid(int n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

bool tr_method_built(int m) {
    return m == kTrV3 || m == kTrV5 || m == kTrSE || m == kTrN1 || m == kTrMUSCLd || m == kTrMUSCLdCRP ||
           m == kTrMC || m == kTrMCCRP || m == kTrWENO || m == kTrWENOCRP;
}

bool tr_method_structured(int m) { return m == kTrWENO || m == kTrWENOCRP; }

hipError_t tr_step(const DevMesh& M, int method, const double* u, double* out, const double* tnx, const double* tny,
                   double* gx, double* gy, const double* fvx, const double* fvy, double dt, int dir, int gnx,
                   int gny, hipStream_t s) {
    const dim3 g = tr_grid(M.n), b(256);
    if (method == kTrWENO || method == kTrWENOCRP) {
        if (method == kTrWENO)
            hipLaunchKernelGGL(k_tr_weno<false>, g, b, 0, s, M, u, fvx, fvy, dt, dir, gnx, gny, out);
        else
            hipLaunchKernelGGL(k_tr_weno<true>, g, b, 0, s, M, u, fvx, fvy, dt, dir, gnx, gny, out);
        return hipGetLastError();
    }
    if (method >= kTrMUSCLd) {
        if (method == kTrMC || method == kTrMCCRP)
            hipLaunchKernelGGL((k_tr_slopes<kLimMC, true>), g, b, 0, s, M, u, gx, gy);
        else
            hipLaunchKernelGGL((k_tr_slopes<kLimMC, false>), g, b, 0, s, M, u, gx, gy);
    }
#define ZFV_TR_FLUX(K) hipLaunchKernelGGL(k_tr_flux<K>, g, b, 0, s, M, u, tnx, tny, gx, gy, fvx, fvy, dt, dir, out)
    switch (method) {
        case kTrV3: ZFV_TR_FLUX(kTrV3); break;
        case kTrV5: ZFV_TR_FLUX(kTrV5); break;
        case kTrSE: ZFV_TR_FLUX(kTrSE); break;
        case kTrN1: ZFV_TR_FLUX(kTrN1); break;
        case kTrMUSCLd: ZFV_TR_FLUX(kTrMUSCLd); break;
        case kTrMUSCLdCRP: ZFV_TR_FLUX(kTrMUSCLdCRP); break;
        case kTrMC: ZFV_TR_FLUX(kTrMC); break;
        default: ZFV_TR_FLUX(kTrMCCRP); break;
    }
#undef ZFV_TR_FLUX
    return hipGetLastError();
}

hipError_t tr_normals(const DevMesh& M, const double* u, double* nx, double* ny, hipStream_t s) {
    hipLaunchKernelGGL(k_tr_normals, tr_grid(M.n), dim3(256), 0, s, M, u, nx, ny);
    return hipGetLastError();
}

hipError_t tr_dt(const DevMesh& M, const double* cvx, const double* cvy, unsigned long long* slot, hipStream_t s) {
    hipLaunchKernelGGL(k_tr_dt, tr_grid(M.n), dim3(256), 0, s, M, cvx, cvy, slot);
    return hipGetLastError();
}

hipError_t tr_flags(const DevMesh& M, const double* u, int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_tr_flags, tr_grid(M.n), dim3(256), 0, s, M, u, flag);
    return hipGetLastError();
}

}  // namespace zfv
