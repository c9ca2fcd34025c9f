// HIP kernels of the single-material finite-volume step for gfx950.
//
// One thread per cell, per-cell gather over the cell's face slots in slot
// order (L, R, B, T, L1, R1, B1, T1): each cell reads its neighbours' states
// and writes only itself, so no atomics and no colouring are needed, and the
// per-cell flux sum has the reference's order (math/solver/sm_fluid.cpp:137-161).
// Slot arrays are slot-major, so a wavefront reads slot s of 64 consecutive
// cells from consecutive addresses.  Cells are mapped to workgroups so that
// each XCD (workgroups are dealt round-robin to the 8 XCDs) owns one
// contiguous range of cells and its stencil neighbours stay in its L2.
//
// The CFL reduction (compute_dt, sm_fluid.cpp:104-114) is fused into the last
// kernel of the previous step: each workgroup reduces min h/(|v|+c) over its
// cells with cross-lane shuffles and LDS, then one 64-bit atomicMin on the bit
// pattern (positive doubles order like unsigned integers).  NaN candidates
// are dropped, as `temp < res` drops them in threads::min (utils/threads.h:292-352).
#include "zfv_internal.hpp"
#include "zfv_step_common.hpp"
#include "zfv_step_faces.hpp"

// Minimum resident waves per SIMD requested from the register allocator for
// the step kernels (caps VGPRs at 512 / waves).
#ifndef ZFV_WAVES_FLUX1
#define ZFV_WAVES_FLUX1 4
#endif
#ifndef ZFV_WAVES_GRAD
#define ZFV_WAVES_GRAD 3
#endif
#ifndef ZFV_WAVES_CORR
#define ZFV_WAVES_CORR 4
#endif
#ifndef ZFV_WAVES_CORRB
#define ZFV_WAVES_CORRB 2
#endif
#ifndef ZFV_WAVES_GRADB
#define ZFV_WAVES_GRADB 2
#endif
#ifndef ZFV_WAVES_FLUX1Q
#define ZFV_WAVES_FLUX1Q 6
#endif
#ifndef ZFV_WAVES_GRADQ
#define ZFV_WAVES_GRADQ 2
#endif
#ifndef ZFV_WAVES_CORRQ
#define ZFV_WAVES_CORRQ 4
#endif

namespace zfv {

namespace {


// Neighbour state through a slot and the LSM offset dr (gradient.h:245-252).
// All loads are unconditional (adj of a boundary slot is the cell itself),
// so a face costs one memory round trip; the slot type then selects.
__device__ __forceinline__ void lsm_neib(const DevMesh& M, const DevState& in, int c, int s,
                                         int t, const Frame& f, const Z& zc, double ccx, double ccy,
                                         Z& zn, double& drx, double& dry) {
    const int so = s * M.stride + c;
    const int nb = M.adj[so];
    const Z zl{in.r[nb], in.u[nb], in.v[nb], in.p[nb], 0.0};
    const double ncx = M.cx[nb], ncy = M.cy[nb];
    const double fx = M.fcx[so], fy = M.fcy[so];
    zn = (t == SC_INNER || t == SC_PERIODIC) ? zl : frame_boundary(f, zc, t == SC_WALL);
    double px = ncx, py = ncy;
    if (t != SC_INNER) frame_symm(f, ccx, ccy, fx, fy, px, py);
    drx = px - ccx;
    dry = py - ccy;
}

}  // namespace

__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long* p, int count, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

__global__ void __launch_bounds__(kMaxBlock) k_dt(DevMesh M, DevState z, Eos eos, unsigned long long* slot) {
    const Tile t = tile_of(M, (int)blockDim.x);
    const int c = t.c0 + (int)threadIdx.x;
    double cand = INFINITY;
    if (c < t.end) cand = cell_dt(load_z(z, c), M.h[c], eos);
    block_min_atomic(slot + 3 * t.pid, cand);
}

// SmFluid::fluxes, first order (sm_fluid.cpp:127-174), plus the next step's
// CFL candidate.
template <int FLUX>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(ZFV_WAVES_FLUX1)))
k_flux1(DevMesh M, DevState in, DevState out, Eos eos, StepCtl ctl) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    const double dt = step_dt(ctl, tl.pid);
    step_prologue(ctl, tl, dt);
    double cand = INFINITY;
    if (c < tl.end) {
        const Z zc = load_z(in, c);
        const int code = M.code[c];
        double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
        bool bad = false;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int nib = (code >> (4 * s)) & 15;
            const int t = nib & 7;
            if (t == SC_UNDEF) continue;
            const Frame fr = slot_frame(M, c, s, nib);
            Z zn;
            if (t <= SC_PERIODIC) zn = load_z(in, M.adj[s * M.stride + c]);
            else zn = frame_boundary(fr, zc, t == SC_WALL);
            Q f = num_flux<FLUX>(frame_local(fr, zc), frame_local(fr, zn), eos, bad);
            f = frame_global(fr, f);
            const double a = M.area[s * M.stride + c];
            F0 = F0 + f.r * a;
            F1 = F1 + f.mx * a;
            F2 = F2 + f.my * a;
            F3 = F3 + f.E * a;
        }
        Q q = to_q(zc);
        const double coef = dt / M.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        axial_source<false>(M, c, zc.p, dt, q);
        const Z zn = to_z(q, eos);
        store_z(out, c, zn);
        if (bad) *ctl.err = 1;
        cand = cell_dt(zn, M.h[c], eos);
    }
    block_min_atomic(&ctl.dtmin[3 * tl.pid + ctl.wr], cand);
}

__device__ __forceinline__ SlotL slot_load(const DevMesh& M, const DevState& in, int c, int s, int nib, const Z& zc,
                                           double ccx, double ccy) {
    SlotL L;
    const Frame fr = slot_frame(M, c, s, nib);
    Z zn;
    lsm_neib(M, in, c, s, nib & 7, fr, zc, ccx, ccy, zn, L.drx, L.dry);
    lsm_drn(fr, L.drx, L.dry, L.dnx, L.dny);
    L.w = M.w[s * M.stride + c];
    L.dF[0] = zn.r - zc.r;
    L.dF[1] = zn.u - zc.u;
    L.dF[2] = zn.v - zc.v;
    L.dF[3] = zn.p - zc.p;
    L.xp = fr.gen || fr.side < 2;
    L.yp = fr.gen || fr.side >= 2;
    return L;
}

// ---- batched loads (the BATCH kernels)
// Four slots s0..s0+3 of a cell are read in two memory round trips: the
// slots' own arrays first, then everything at the neighbours.  No load sits
// under a condition: an undefined slot (adj -1, or anything outside the
// mesh) reads the cell itself, and the slot types only select among the
// loaded values afterwards, so all of a cell's loads of one round are in
// flight together.  The general normals are read only when one of the four
// slots has one (canonical normals are not stored per slot).
__device__ __forceinline__ int slot_nb(const DevMesh& M, int c, int a) {
    return (unsigned)a < (unsigned)M.n ? a : c;
}

__device__ __forceinline__ bool any_general(int code, int s0) {
    return ((code >> (4 * s0)) & 0x8888) != 0;
}

struct LsmRaw4 {
    double fx[4], fy[4], w[4], a[4], nx[4], ny[4];
    double r[4], u[4], v[4], p[4], cx[4], cy[4];
};

template <bool AREA>
__device__ __forceinline__ void lsm_load4(const DevMesh& M, const DevState& in, int c, int s0, int code,
                                          LsmRaw4& R) {
    int adj[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int so = (s0 + k) * M.stride + c;
        adj[k] = M.adj[so];
        R.fx[k] = M.fcx[so];
        R.fy[k] = M.fcy[so];
        R.w[k] = M.w[so];
        R.a[k] = AREA ? M.area[so] : 0.0;
        R.nx[k] = 0.0;
        R.ny[k] = 0.0;
    }
    if (any_general(code, s0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            R.nx[k] = M.nx[(s0 + k) * M.stride + c];
            R.ny[k] = M.ny[(s0 + k) * M.stride + c];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int nb = slot_nb(M, c, adj[k]);
        R.r[k] = in.r[nb];
        R.u[k] = in.u[nb];
        R.v[k] = in.v[nb];
        R.p[k] = in.p[nb];
        R.cx[k] = M.cx[nb];
        R.cy[k] = M.cy[nb];
    }
}

// slot_load's values from the loads of slot s = s0 + k: the same operations
// in the same order (lsm_neib, lsm_drn).
__device__ __forceinline__ SlotL slot_eval(const LsmRaw4& R, int k, int s, int nib, const Z& zc, double ccx,
                                           double ccy) {
    SlotL L;
    const bool gen = (nib & kGeneralNormal) != 0;
    const Frame fr = frame_of(s & 3, gen, gen ? R.nx[k] : 0.0, gen ? R.ny[k] : 0.0);
    const int t = nib & 7;
    const Z zl{R.r[k], R.u[k], R.v[k], R.p[k], 0.0};
    const Z zn = (t == SC_INNER || t == SC_PERIODIC) ? zl : frame_boundary(fr, zc, t == SC_WALL);
    double px = R.cx[k], py = R.cy[k];
    if (t != SC_INNER) frame_symm(fr, ccx, ccy, R.fx[k], R.fy[k], px, py);
    L.drx = px - ccx;
    L.dry = py - ccy;
    lsm_drn(fr, L.drx, L.dry, L.dnx, L.dny);
    L.w = R.w[k];
    L.dF[0] = zn.r - zc.r;
    L.dF[1] = zn.u - zc.u;
    L.dF[2] = zn.v - zc.v;
    L.dF[3] = zn.p - zc.p;
    L.xp = fr.gen || fr.side < 2;
    L.yp = fr.gen || fr.side >= 2;
    return L;
}

// Predictor face geometry of four slots (face centre, area, normal), one
// round trip.
struct PredRaw4 {
    double fx[4], fy[4], a[4], nx[4], ny[4];
};

__device__ __forceinline__ void pred_load4(const DevMesh& M, int c, int s0, int code, PredRaw4& P) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int so = (s0 + k) * M.stride + c;
        P.fx[k] = M.fcx[so];
        P.fy[k] = M.fcy[so];
        P.a[k] = M.area[so];
        P.nx[k] = 0.0;
        P.ny[k] = 0.0;
    }
    if (any_general(code, s0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            P.nx[k] = M.nx[(s0 + k) * M.stride + c];
            P.ny[k] = M.ny[(s0 + k) * M.stride + c];
        }
    }
}

__device__ __forceinline__ Frame raw_frame(int s, int nib, double nx, double ny) {
    const bool gen = (nib & kGeneralNormal) != 0;
    return frame_of(s & 3, gen, gen ? nx : 0.0, gen ? ny : 0.0);
}

// Everything four corrector faces read (corr_load of slots s0..s0+3), with
// the face normals as stored: the frames are made when a face is evaluated
// (face_of), which keeps the registers of four faces' loads within what two
// resident waves per SIMD allow.
struct FaceRaw {
    Z hn;
    double ngx[4], ngy[4];
    double ncx, ncy, fx, fy, a, nx, ny;
    int wnb;
};

__device__ __forceinline__ void corr_load4(const DevMesh& M, const DevState& half, const DevGrad& G,
                                           const int* __restrict__ wait_in, int c, int s0, int code, FaceRaw* F) {
    int adj[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int so = (s0 + k) * M.stride + c;
        adj[k] = M.adj[so];
        F[k].fx = M.fcx[so];
        F[k].fy = M.fcy[so];
        F[k].a = M.area[so];
        F[k].nx = 0.0;
        F[k].ny = 0.0;
    }
    if (any_general(code, s0)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            F[k].nx = M.nx[(s0 + k) * M.stride + c];
            F[k].ny = M.ny[(s0 + k) * M.stride + c];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int nb = slot_nb(M, c, adj[k]);
        F[k].hn = load_z(half, nb);
        F[k].wnb = wait_in[nb];
        F[k].ncx = M.cx[nb];
        F[k].ncy = M.cy[nb];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            F[k].ngx[j] = G.dx[j * M.stride + nb];
            F[k].ngy[j] = G.dy[j * M.stride + nb];
        }
    }
}

__device__ __forceinline__ FaceC face_of(const FaceRaw& R, int s, int nib) {
    FaceC F;
    F.hn = R.hn;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        F.ngx[j] = R.ngx[j];
        F.ngy[j] = R.ngy[j];
    }
    F.ncx = R.ncx;
    F.ncy = R.ncy;
    F.fx = R.fx;
    F.fy = R.fy;
    F.a = R.a;
    F.fr = raw_frame(s, nib, R.nx, R.ny);
    F.wnb = R.wnb;
    return F;
}

__device__ __forceinline__ Q pred_face(const DevMesh& M, int c, int s, int nib, const Z& zc, double ccx, double ccy,
                                       const double* gx, const double* gy, const Eos& eos) {
    const int so = s * M.stride + c;
    return pred_eval(slot_frame(M, c, s, nib), M.fcx[so] - ccx, M.fcy[so] - ccy, M.area[so], zc, gx, gy, eos);
}

// SmFluid::compute_grad (sm_fluid.cpp:116-125): gradient::LSM
// (gradient.h:226-282) then gradient::limiting (gradient.h:412-474) on
// {rho, u, v, p}, fused with SmFluid::fluxes_stage1 (sm_fluid.cpp:176-253),
// which only reads the cell's own state and freshly limited gradient.
//
// For a canonical normal drn = (dr.n) n keeps one component of dr, so the
// 3x3 LSM matrix has the 2x2 block A = sum w drn dr^T, A(2,2) = 1, and
// Eigen's cofactor inverse reduces to adj(A) * (1/det) with
// det = A11*A00 + (-A01)*A10.
// Without the predictor, `raw` (if not null) receives the unlimited LSM
// gradient of rho ([2][stride]) for the refinement criterion.  BATCH (the
// batched layout): the four main slots' loads, the predictor's face
// geometry included, are issued together and unconditionally (lsm_load4):
// two memory round trips per cell instead of two per slot, and two more per
// pass for the sub-face slots of a cell with hanging faces.
template <int LIM, bool PRED, bool BATCH>
__global__ void __launch_bounds__(kMaxBlock)
__attribute__((amdgpu_waves_per_eu(BATCH ? ZFV_WAVES_GRADB : ZFV_WAVES_GRAD)))
k_grad_pred(DevMesh M, DevState in, DevGrad G, DevState half,
                                                      const int* __restrict__ wait, Eos eos, StepCtl ctl,
                                                      double* __restrict__ raw) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    const double dt = step_dt(ctl, tl.pid);
    if (PRED) {
        if (tl.first && threadIdx.x == 0) ctl.dtmin[3 * tl.pid + ctl.clr] = kInfBits;
    }
    if (c >= tl.end) return;

    const Z zc = load_z(in, c);
    const double ccx = M.cx[c], ccy = M.cy[c];
    const int code = M.code[c];
    const int wc = PRED ? wait[c] : 0;
    const bool sub = (code >> 16) != 0;   // some second sub-face is defined

    // The four main slots are loaded once and kept for the limiting pass;
    // sub-face slots (hanging faces) follow in slot order.  BATCH: two
    // round trips for the four main slots (lsm_load4), two for the four
    // sub-face slots of a cell that has any.
    SlotL S[4];
    double pdx[4], pdy[4], pa[4];
    if (BATCH) {
        LsmRaw4 R;
        lsm_load4<PRED>(M, in, c, 0, code, R);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int nib = (code >> (4 * s)) & 15;
            const bool def = (nib & 7) != SC_UNDEF;
            S[s] = def ? slot_eval(R, s, s, nib, zc, ccx, ccy) : SlotL{};
            pdx[s] = def ? R.fx[s] - ccx : 0.0;
            pdy[s] = def ? R.fy[s] - ccy : 0.0;
            pa[s] = def ? R.a[s] : 0.0;
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int nib = (code >> (4 * s)) & 15;
            S[s] = (nib & 7) != SC_UNDEF ? slot_load(M, in, c, s, nib, zc, ccx, ccy) : SlotL{};
        }
    }
    double A00 = 0.0, A01 = 0.0, A10 = 0.0, A11 = 0.0;
    double Fx[4] = {0.0, 0.0, 0.0, 0.0}, Fy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (((code >> (4 * s)) & 7) != SC_UNDEF) lsm_add(S[s], A00, A01, A10, A11, Fx, Fy);
    if (sub) {
        if (BATCH) {
            LsmRaw4 R;
            lsm_load4<false>(M, in, c, 4, code, R);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int nib = (code >> (4 * (4 + k))) & 15;
                if ((nib & 7) == SC_UNDEF) continue;
                lsm_add(slot_eval(R, k, 4 + k, nib, zc, ccx, ccy), A00, A01, A10, A11, Fx, Fy);
            }
        } else {
            for (int s = 4; s < 8; ++s) {
                const int nib = (code >> (4 * s)) & 15;
                if ((nib & 7) == SC_UNDEF) continue;
                lsm_add(slot_load(M, in, c, s, nib, zc, ccx, ccy), A00, A01, A10, A11, Fx, Fy);
            }
        }
    }
    if (fabs(A11) < 1e-8) A11 = 1.0;   // A(2,2) = 0 -> 1 always holds in 2D (gradient.h:268-273)
    const double det = A11 * A00 + (-A01) * A10;
    const double invdet = 1.0 / det;
    const double B00 = A11 * invdet, B01 = (-A01) * invdet;
    const double B10 = (-A10) * invdet, B11 = A00 * invdet;
    double gx[4], gy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gx[k] = B00 * Fx[k] + B01 * Fy[k];
        gy[k] = B10 * Fx[k] + B11 * Fy[k];
    }
    if (!PRED && raw) {
        raw[c] = gx[0];
        raw[M.stride + c] = gy[0];
    }

    // limiting: limiter(2 du_n - dF_p, dF_p) * dF_p re-solved with the same A
    double Lx[4] = {0.0, 0.0, 0.0, 0.0}, Ly[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (((code >> (4 * s)) & 7) != SC_UNDEF) lim_add<LIM>(S[s], gx, gy, Lx, Ly);
    if (sub) {
        if (BATCH) {
            LsmRaw4 R;
            lsm_load4<false>(M, in, c, 4, code, R);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int nib = (code >> (4 * (4 + k))) & 15;
                if ((nib & 7) == SC_UNDEF) continue;
                lim_add<LIM>(slot_eval(R, k, 4 + k, nib, zc, ccx, ccy), gx, gy, Lx, Ly);
            }
        } else {
            for (int s = 4; s < 8; ++s) {
                const int nib = (code >> (4 * s)) & 15;
                if ((nib & 7) == SC_UNDEF) continue;
                lim_add<LIM>(slot_load(M, in, c, s, nib, zc, ccx, ccy), gx, gy, Lx, Ly);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gx[k] = B00 * Lx[k] + B01 * Ly[k];
        gy[k] = B10 * Lx[k] + B11 * Ly[k];
        G.dx[k * M.stride + c] = gx[k];
        G.dy[k * M.stride + c] = gy[k];
    }
    if (!PRED) return;

    // predictor: half = PState(q - (dt/2V) sum Flux(z_c + grad.(x_f - x_c)) S)
    Z zh = zc;
    if (wc <= 0) {
        double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
        if (BATCH) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int nib = (code >> (4 * s)) & 15;
                if ((nib & 7) == SC_UNDEF) continue;
                const Q f = pred_eval(slot_frame(M, c, s, nib), pdx[s], pdy[s], pa[s], zc, gx, gy, eos);
                F0 = F0 + f.r;
                F1 = F1 + f.mx;
                F2 = F2 + f.my;
                F3 = F3 + f.E;
            }
            if (sub) {
                PredRaw4 P;
                pred_load4(M, c, 4, code, P);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int nib = (code >> (4 * (4 + k))) & 15;
                    if ((nib & 7) == SC_UNDEF) continue;
                    const Q f = pred_eval(raw_frame(4 + k, nib, P.nx[k], P.ny[k]), P.fx[k] - ccx, P.fy[k] - ccy,
                                          P.a[k], zc, gx, gy, eos);
                    F0 = F0 + f.r;
                    F1 = F1 + f.mx;
                    F2 = F2 + f.my;
                    F3 = F3 + f.E;
                }
            }
        } else {
            for (int s = 0; s < 8; ++s) {
                const int nib = (code >> (4 * s)) & 15;
                if ((nib & 7) == SC_UNDEF) continue;
                const Q f = pred_face(M, c, s, nib, zc, ccx, ccy, gx, gy, eos);
                F0 = F0 + f.r;
                F1 = F1 + f.mx;
                F2 = F2 + f.my;
                F3 = F3 + f.E;
            }
        }
        Q q = to_q(zc);
        const double coef = 0.5 * dt / M.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        axial_source<true>(M, c, zc.p, dt, q);
        zh = to_z(q, eos);
        if (is_bad(zh, eos)) zh = zc;
    }
    store_z(half, c, zh);
}

__device__ __forceinline__ FaceC corr_load(const DevMesh& M, const DevState& half, const DevGrad& G,
                                           const int* __restrict__ wait_in, int c, int s, int nib) {
    FaceC F;
    const int so = s * M.stride + c;
    const int nb = M.adj[so];
    F.fx = M.fcx[so];
    F.fy = M.fcy[so];
    F.a = M.area[so];
    F.hn = load_z(half, nb);
    F.wnb = wait_in[nb];
    F.ncx = M.cx[nb];
    F.ncy = M.cy[nb];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        F.ngx[k] = G.dx[k * M.stride + nb];
        F.ngy[k] = G.dy[k * M.stride + nb];
    }
    F.fr = slot_frame(M, c, s, nib);
    return F;
}

template <int FLUX>
__device__ __forceinline__ Q corr_face(const DevMesh& M, const DevState& half, const DevGrad& G,
                                       const int* __restrict__ wait_in, int c, int s, int nib, const Z& zch, int wc,
                                       double ccx, double ccy, const double* gx, const double* gy, const Eos& eos,
                                       bool& bad) {
    return corr_eval<FLUX>(corr_load(M, half, G, wait_in, c, s, nib), nib, zch, wc, ccx, ccy, gx, gy, eos, bad);
}

// SmFluid::fluxes_stage2, corrector (sm_fluid.cpp:255-366), plus the next
// step's CFL candidate.  `wait` is double-buffered: neighbours' isolation is
// read from the pre-pass array.  BATCH (the batched layout): the loads of
// the four main slots are issued together, unconditionally, before any face
// is evaluated (corr_load4), two dependent memory round trips per cell
// instead of two per face, and two more for the four sub-face slots of a
// cell with hanging faces, at twice the registers (fewer resident waves; for
// meshes too small to fill the GPU); otherwise one face at a time.  The flux
// terms are summed in slot order either way.
template <int FLUX, bool BATCH>
__global__ void __launch_bounds__(kMaxBlock)
__attribute__((amdgpu_waves_per_eu(BATCH ? ZFV_WAVES_CORRB : ZFV_WAVES_CORR)))
k_corr(DevMesh M, DevState init, DevState half, DevGrad G, const int* __restrict__ wait_in,
       int* __restrict__ wait_out, DevState next, Eos eos, StepCtl ctl) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    const double dt = step_dt(ctl, tl.pid);
    if (tl.first && threadIdx.x == 0) ctl.dt_out[tl.pid] = dt;
    double cand = INFINITY;
    if (c < tl.end) {
        const Z zch = load_z(half, c);
        const double ccx = M.cx[c], ccy = M.cy[c];
        const int code = M.code[c];
        const int wc = wait_in[c];
        double gx[4], gy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gx[k] = G.dx[k * M.stride + c];
            gy[k] = G.dy[k * M.stride + c];
        }
        double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
        bool bad = false;
        if (BATCH) {
            FaceRaw D[4];
            corr_load4(M, half, G, wait_in, c, 0, code, D);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int nib = (code >> (4 * s)) & 15;
                if ((nib & 7) == SC_UNDEF) continue;
                const Q f = corr_eval<FLUX>(face_of(D[s], s, nib), nib, zch, wc, ccx, ccy, gx, gy, eos, bad);
                F0 = F0 + f.r;
                F1 = F1 + f.mx;
                F2 = F2 + f.my;
                F3 = F3 + f.E;
            }
            if ((code >> 16) != 0) {
                corr_load4(M, half, G, wait_in, c, 4, code, D);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int nib = (code >> (4 * (4 + k))) & 15;
                    if ((nib & 7) == SC_UNDEF) continue;
                    const Q f = corr_eval<FLUX>(face_of(D[k], 4 + k, nib), nib, zch, wc, ccx, ccy, gx, gy, eos, bad);
                    F0 = F0 + f.r;
                    F1 = F1 + f.mx;
                    F2 = F2 + f.my;
                    F3 = F3 + f.E;
                }
            }
        } else {
            for (int s = 0; s < 8; ++s) {
                const int nib = (code >> (4 * s)) & 15;
                if ((nib & 7) == SC_UNDEF) continue;
                const Q f = corr_face<FLUX>(M, half, G, wait_in, c, s, nib, zch, wc, ccx, ccy, gx, gy, eos, bad);
                F0 = F0 + f.r;
                F1 = F1 + f.mx;
                F2 = F2 + f.my;
                F3 = F3 + f.E;
            }
        }
        const Z zc = load_z(init, c);
        Q q = to_q(zc);
        const double coef = dt / M.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        axial_source<false>(M, c, zch.p, dt, q);
        Z zn = to_z(q, eos);
        int wn;
        if (is_bad(zn, eos)) {
            zn = zc;
            wn = 3;
            atomicOr(ctl.err, kErrIsolated);
        } else {
            wn = wc > 0 ? wc - 1 : wc;
        }
        store_z(next, c, zn);
        wait_out[c] = wn;
        if (bad) *ctl.err = 1;
        cand = cell_dt(zn, M.h[c], eos);
    }
    block_min_atomic(&ctl.dtmin[3 * tl.pid + ctl.wr], cand);
}

// ---------------------------------------------------------------- quad layout
// Four lanes per cell (a lane quad): lane j of the quad evaluates slot j
// (side L, R, B, T) and, when it is defined, slot j + 4.  The per-slot terms
// are combined in slot order through DPP quad broadcasts, so every sum is the
// same sequence of additions as in the one-thread-per-cell kernels above and
// the results are bit-identical; a wavefront holds 16 cells and each lane
// carries one face, which gives four times the waves per cell count.
namespace {

template <int J>
__device__ __forceinline__ double qget(double v) {
    constexpr int ctrl = J * 0x55;   // quad_perm [J, J, J, J]
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sums in slot order.  qlo: 0.0 plus, for each main slot s = 0..3 set in
// `mask`, the term held by lane s of the quad.  qhi: continues `acc` with
// the sub-slots s = 4..7 set in `mask`, the term of slot s held by lane s - 4.
// qlo then qhi is the sequential slot-order sum over all eight slots.
__device__ __forceinline__ double qlo(double v, unsigned mask) {
    double acc = 0.0, x;
    x = qget<0>(v); if (mask & 1u) acc = acc + x;
    x = qget<1>(v); if (mask & 2u) acc = acc + x;
    x = qget<2>(v); if (mask & 4u) acc = acc + x;
    x = qget<3>(v); if (mask & 8u) acc = acc + x;
    return acc;
}
__device__ __forceinline__ double qhi(double acc, double v, unsigned mask) {
    double x;
    x = qget<0>(v); if (mask & 16u) acc = acc + x;
    x = qget<1>(v); if (mask & 32u) acc = acc + x;
    x = qget<2>(v); if (mask & 64u) acc = acc + x;
    x = qget<3>(v); if (mask & 128u) acc = acc + x;
    return acc;
}



// Slot masks of a code word: defined slots, and the slots whose LSM x / y
// rows receive a term (canonical L/R normals have drn.y = 0, B/T drn.x = 0).
struct Masks {
    unsigned def, xm, ym;
};
__device__ __forceinline__ Masks slot_masks(int code) {
    unsigned def = 0u, gen = 0u;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int nib = (code >> (4 * s)) & 15;
        if (nib & 7) def |= 1u << s;
        if (nib & kGeneralNormal) gen |= 1u << s;
    }
    return Masks{def, def & (gen | 0x33u), def & (gen | 0xCCu)};
}

template <class T>
__device__ __forceinline__ T by_side(int side, T a, T b, T c, T d) {
    return side == 0 ? a : (side == 1 ? b : (side == 2 ? c : d));
}

// Store a state: lane j of the quad writes rho / u / v / p, lane 0 also e.
__device__ __forceinline__ void store_z_quad(const DevState& s, int c, int side, const Z& z) {
    double* dst = by_side(side, s.r, s.u, s.v, s.p);
    dst[c] = by_side(side, z.r, z.u, z.v, z.p);
    if (side == 0) s.e[c] = z.e;
}

// Flux term of slot s (f * area), first order.
template <int FLUX>
__device__ __forceinline__ Q flux1_term(const DevMesh& M, const DevState& in, const Z& zc, int c, int s, int nib,
                                        const Eos& eos, bool& bad) {
    const int t = nib & 7;
    const Frame fr = slot_frame(M, c, s, nib);
    Z zn;
    if (t <= SC_PERIODIC) zn = load_z(in, M.adj[s * M.stride + c]);
    else zn = frame_boundary(fr, zc, t == SC_WALL);
    Q f = num_flux<FLUX>(frame_local(fr, zc), frame_local(fr, zn), eos, bad);
    f = frame_global(fr, f);
    const double a = M.area[s * M.stride + c];
    return Q{f.r * a, f.mx * a, f.my * a, f.E * a};
}

// Slot-order sums of the four flux components: main slots from t0, then
// (SUB) the sub-slots from t1.
template <bool SUB>
__device__ __forceinline__ void qsum4(const Q& t0, const Q& t1, unsigned mask, double& F0, double& F1, double& F2,
                                      double& F3) {
    F0 = qlo(t0.r, mask);
    F1 = qlo(t0.mx, mask);
    F2 = qlo(t0.my, mask);
    F3 = qlo(t0.E, mask);
    if (SUB) {
        F0 = qhi(F0, t1.r, mask);
        F1 = qhi(F1, t1.mx, mask);
        F2 = qhi(F2, t1.my, mask);
        F3 = qhi(F3, t1.E, mask);
    }
}

}  // namespace

template <int FLUX>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(ZFV_WAVES_FLUX1Q)))
k_flux1_q(DevMesh M, DevState in, DevState out, Eos eos, StepCtl ctl) {
    const Tile tl = tile_of(M, (int)blockDim.x >> 2);
    const int c = tl.c0 + ((int)threadIdx.x >> 2), side = threadIdx.x & 3;
    const double dt = step_dt(ctl, tl.pid);
    step_prologue(ctl, tl, dt);
    double cand = INFINITY;
    const bool valid = c < tl.end;
    const int code = valid ? M.code[c] : 0;
    const bool sub = __any((code >> 16) != 0);
    if (valid) {
        const Z zc = load_z(in, c);
        const Masks mk = slot_masks(code);
        bool bad = false;
        Q t0{0.0, 0.0, 0.0, 0.0}, t1{0.0, 0.0, 0.0, 0.0};
        const int nib0 = (code >> (4 * side)) & 15;
        if (nib0 & 7) t0 = flux1_term<FLUX>(M, in, zc, c, side, nib0, eos, bad);
        const int nib1 = (code >> (4 * (side + 4))) & 15;
        if (nib1 & 7) t1 = flux1_term<FLUX>(M, in, zc, c, side + 4, nib1, eos, bad);
        double F0, F1, F2, F3;
        if (sub) qsum4<true>(t0, t1, mk.def, F0, F1, F2, F3);
        else qsum4<false>(t0, t1, mk.def, F0, F1, F2, F3);
        Q q = to_q(zc);
        const double coef = dt / M.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        axial_source<false>(M, c, zc.p, dt, q);
        const Z zn = to_z(q, eos);
        store_z_quad(out, c, side, zn);
        if (bad) *ctl.err = 1;
        cand = cell_dt(zn, M.h[c], eos);
    }
    block_min_atomic(&ctl.dtmin[3 * tl.pid + ctl.wr], cand);
}

namespace {

// LSM data of one slot (gradient.h:245-256).
struct LsmSlot {
    Z zn;
    double drx, dry, dnx, dny, w;
};

__device__ __forceinline__ LsmSlot lsm_slot(const DevMesh& M, const DevState& in, int c, int s, int nib,
                                            const Z& zc, double ccx, double ccy) {
    LsmSlot L;
    const Frame fr = slot_frame(M, c, s, nib);
    lsm_neib(M, in, c, s, nib & 7, fr, zc, ccx, ccy, L.zn, L.drx, L.dry);
    lsm_drn(fr, L.drx, L.dry, L.dnx, L.dny);
    L.w = M.w[s * M.stride + c];
    return L;
}

// Predictor flux term of slot s (sm_fluid.cpp:176-253).
__device__ __forceinline__ Q pred_term(const DevMesh& M, int c, int s, int nib, const Z& zc, double ccx, double ccy,
                                       const double* gx, const double* gy, const Eos& eos) {
    const Frame fr = slot_frame(M, c, s, nib);
    const double dx = M.fcx[s * M.stride + c] - ccx;
    const double dy = M.fcy[s * M.stride + c] - ccy;
    Z zm = extrap(zc, gx, gy, dx, dy);
    zm.e = eos_e(eos, zm.r, zm.p);
    if (is_bad(zm, eos)) zm = zc;
    Q f = phys_flux(frame_local(fr, zm));
    f = frame_global(fr, f);
    const double a = M.area[s * M.stride + c];
    return Q{f.r * a, f.mx * a, f.my * a, f.E * a};
}

}  // namespace

// LSM pass-1 terms of one slot added to the running slot-order sums.
template <bool HI>
__device__ __forceinline__ void lsm_acc(const LsmSlot& L, const Z& zc, const Masks& mk, double* A, double* Fx,
                                        double* Fy) {
    const double wx = L.w * L.dnx, wy = L.w * L.dny;
    const double d[4] = {L.zn.r - zc.r, L.zn.u - zc.u, L.zn.v - zc.v, L.zn.p - zc.p};
    if (HI) {
        A[0] = qhi(A[0], wx * L.drx, mk.xm);
        A[1] = qhi(A[1], wx * L.dry, mk.xm);
        A[2] = qhi(A[2], wy * L.drx, mk.ym);
        A[3] = qhi(A[3], wy * L.dry, mk.ym);
    } else {
        A[0] = qlo(wx * L.drx, mk.xm);
        A[1] = qlo(wx * L.dry, mk.xm);
        A[2] = qlo(wy * L.drx, mk.ym);
        A[3] = qlo(wy * L.dry, mk.ym);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double e = L.w * d[k];
        if (HI) {
            Fx[k] = qhi(Fx[k], e * L.dnx, mk.xm);
            Fy[k] = qhi(Fy[k], e * L.dny, mk.ym);
        } else {
            Fx[k] = qlo(e * L.dnx, mk.xm);
            Fy[k] = qlo(e * L.dny, mk.ym);
        }
    }
}

// Limiting pass terms of one slot (gradient.h:412-474) added to the sums.
template <int LIM, bool HI>
__device__ __forceinline__ void lim_acc(const LsmSlot& L, const Z& zc, const Masks& mk, const double* gx,
                                        const double* gy, double* Lx, double* Ly) {
    const double d[4] = {L.zn.r - zc.r, L.zn.u - zc.u, L.zn.v - zc.v, L.zn.p - zc.p};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double du_n = gx[k] * L.drx + gy[k] * L.dry;
        const double lim = limiter2<LIM>(2.0 * du_n - d[k], d[k]) * d[k];
        const double e = L.w * lim;
        if (HI) {
            Lx[k] = qhi(Lx[k], e * L.dnx, mk.xm);
            Ly[k] = qhi(Ly[k], e * L.dny, mk.ym);
        } else {
            Lx[k] = qlo(e * L.dnx, mk.xm);
            Ly[k] = qlo(e * L.dny, mk.ym);
        }
    }
}

template <int LIM, bool PRED, bool SUB>
__device__ __forceinline__ void grad_pred_quad(const DevMesh& M, const DevState& in, const DevGrad& G,
                                               const DevState& half, const int* __restrict__ wait, const Eos& eos,
                                               double dt, double* __restrict__ raw, int c, int side, int code) {
    const Z zc = load_z(in, c);
    const double ccx = M.cx[c], ccy = M.cy[c];
    const Masks mk = slot_masks(code);
    const int nib0 = (code >> (4 * side)) & 15;
    const int nib1 = (code >> (4 * (side + 4))) & 15;
    LsmSlot L0{};
    if (nib0 & 7) L0 = lsm_slot(M, in, c, side, nib0, zc, ccx, ccy);

    // pass 1: A = sum w drn dr^T, F = sum w dU drn, in slot order
    double A[4], Fx[4], Fy[4];
    lsm_acc<false>(L0, zc, mk, A, Fx, Fy);
    if (SUB) {
        LsmSlot L1{};
        if (nib1 & 7) L1 = lsm_slot(M, in, c, side + 4, nib1, zc, ccx, ccy);
        lsm_acc<true>(L1, zc, mk, A, Fx, Fy);
    }
    const double A00 = A[0], A01 = A[1], A10 = A[2];
    double A11 = A[3];
    if (fabs(A11) < 1e-8) A11 = 1.0;   // A(2,2) = 0 -> 1 always holds in 2D (gradient.h:268-273)
    const double det = A11 * A00 + (-A01) * A10;
    const double invdet = 1.0 / det;
    const double B00 = A11 * invdet, B01 = (-A01) * invdet;
    const double B10 = (-A10) * invdet, B11 = A00 * invdet;
    double gx[4], gy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gx[k] = B00 * Fx[k] + B01 * Fy[k];
        gy[k] = B10 * Fx[k] + B11 * Fy[k];
    }
    if (!PRED && raw && side == 0) {
        raw[c] = gx[0];
        raw[M.stride + c] = gy[0];
    }

    // pass 2: limited differences, re-solved with the same A
    double Lx[4], Ly[4];
    lim_acc<LIM, false>(L0, zc, mk, gx, gy, Lx, Ly);
    if (SUB) {
        LsmSlot L1{};
        if (nib1 & 7) L1 = lsm_slot(M, in, c, side + 4, nib1, zc, ccx, ccy);
        lim_acc<LIM, true>(L1, zc, mk, gx, gy, Lx, Ly);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gx[k] = B00 * Lx[k] + B01 * Ly[k];
        gy[k] = B10 * Lx[k] + B11 * Ly[k];
    }
    G.dx[side * M.stride + c] = by_side(side, gx[0], gx[1], gx[2], gx[3]);
    G.dy[side * M.stride + c] = by_side(side, gy[0], gy[1], gy[2], gy[3]);
    if (!PRED) return;

    // predictor: half = PState(q - (dt/2V) sum Flux(z_c + grad.(x_f - x_c)) S)
    Z zh = zc;
    if (wait[c] <= 0) {
        Q t0{0.0, 0.0, 0.0, 0.0}, t1{0.0, 0.0, 0.0, 0.0};
        if (nib0 & 7) t0 = pred_term(M, c, side, nib0, zc, ccx, ccy, gx, gy, eos);
        if (SUB && (nib1 & 7)) t1 = pred_term(M, c, side + 4, nib1, zc, ccx, ccy, gx, gy, eos);
        double F0, F1, F2, F3;
        qsum4<SUB>(t0, t1, mk.def, F0, F1, F2, F3);
        Q q = to_q(zc);
        const double coef = 0.5 * dt / M.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        axial_source<true>(M, c, zc.p, dt, q);
        zh = to_z(q, eos);
        if (is_bad(zh, eos)) zh = zc;
    }
    store_z_quad(half, c, side, zh);
}

template <int LIM, bool PRED>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(ZFV_WAVES_GRADQ)))
k_grad_pred_q(DevMesh M, DevState in, DevGrad G, DevState half, const int* __restrict__ wait, Eos eos, StepCtl ctl,
              double* __restrict__ raw) {
    const Tile tl = tile_of(M, (int)blockDim.x >> 2);
    const int c = tl.c0 + ((int)threadIdx.x >> 2), side = threadIdx.x & 3;
    const double dt = step_dt(ctl, tl.pid);
    if (PRED) {
        if (tl.first && threadIdx.x == 0) ctl.dtmin[3 * tl.pid + ctl.clr] = kInfBits;
    }
    const bool valid = c < tl.end;
    const int code = valid ? M.code[c] : 0;
    const bool sub = __any((code >> 16) != 0);
    if (!valid) return;
    if (sub) grad_pred_quad<LIM, PRED, true>(M, in, G, half, wait, eos, dt, raw, c, side, code);
    else grad_pred_quad<LIM, PRED, false>(M, in, G, half, wait, eos, dt, raw, c, side, code);
}

namespace {

}  // namespace

template <int FLUX>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(ZFV_WAVES_CORRQ)))
k_corr_q(DevMesh M, DevState init, DevState half, DevGrad G, const int* __restrict__ wait_in,
         int* __restrict__ wait_out, DevState next, Eos eos, StepCtl ctl) {
    const Tile tl = tile_of(M, (int)blockDim.x >> 2);
    const int c = tl.c0 + ((int)threadIdx.x >> 2), side = threadIdx.x & 3;
    const double dt = step_dt(ctl, tl.pid);
    if (tl.first && threadIdx.x == 0) ctl.dt_out[tl.pid] = dt;
    double cand = INFINITY;
    const bool valid = c < tl.end;
    const int code = valid ? M.code[c] : 0;
    const bool sub = __any((code >> 16) != 0);
    if (valid) {
        const Z zch = load_z(half, c);
        const double ccx = M.cx[c], ccy = M.cy[c];
        const int wc = wait_in[c];
        double gx[4], gy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gx[k] = G.dx[k * M.stride + c];
            gy[k] = G.dy[k * M.stride + c];
        }
        const Masks mk = slot_masks(code);
        bool bad = false;
        Q t0{0.0, 0.0, 0.0, 0.0}, t1{0.0, 0.0, 0.0, 0.0};
        const int nib0 = (code >> (4 * side)) & 15;
        if (nib0 & 7) t0 = corr_face<FLUX>(M, half, G, wait_in, c, side, nib0, zch, wc, ccx, ccy, gx, gy, eos, bad);
        const int nib1 = (code >> (4 * (side + 4))) & 15;
        if (nib1 & 7)
            t1 = corr_face<FLUX>(M, half, G, wait_in, c, side + 4, nib1, zch, wc, ccx, ccy, gx, gy, eos, bad);
        double F0, F1, F2, F3;
        if (sub) qsum4<true>(t0, t1, mk.def, F0, F1, F2, F3);
        else qsum4<false>(t0, t1, mk.def, F0, F1, F2, F3);
        const Z zc = load_z(init, c);
        Q q = to_q(zc);
        const double coef = dt / M.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        axial_source<false>(M, c, zch.p, dt, q);
        Z zn = to_z(q, eos);
        int wn;
        if (is_bad(zn, eos)) {
            zn = zc;
            wn = 3;
            atomicOr(ctl.err, kErrIsolated);
        } else {
            wn = wc > 0 ? wc - 1 : wc;
        }
        store_z_quad(next, c, side, zn);
        if (side == 0) wait_out[c] = wn;
        if (bad) *ctl.err = 1;
        cand = cell_dt(zn, M.h[c], eos);
    }
    block_min_atomic(&ctl.dtmin[3 * tl.pid + ctl.wr], cand);
}

// Per-workgroup partial sums of V*q (conserved totals).
__global__ void __launch_bounds__(kBlock) k_totals(DevMesh M, DevState z, double* partial) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < M.n) {
        const Q q = to_q(load_z(z, c));
        const double V = M.vol[c];
        v[0] = V * q.r;
        v[1] = V * q.mx;
        v[2] = V * q.my;
        v[3] = V * q.E;
    }
    __shared__ double red[4][kBlock];
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------- scratch
// Device memory of the entry points that allocate nothing
// (zfv_conserved_totals_noalloc): [0, 4) the result, then the per-workgroup
// partial sums.  One copy per device, held until the process ends and shared
// by every call on that device (calls that overlap on one device race).  In
// host code `scratch` names a host-side stand-in, not device memory: the
// host uses only the address scratch_address() returns.
static __device__ __align__(256) double scratch[kScratchDoubles];

// The four totals, each summed over the workgroups' partials in workgroup
// order from 0.0: the additions zfv_conserved_totals makes on the host.
__global__ void __launch_bounds__(64) k_totals_sum(const double* __restrict__ partial, int nblocks,
                                                   double* __restrict__ out) {
    const int k = (int)threadIdx.x;
    if (k >= 4) return;
    double acc = 0.0;
    for (int b = 0; b < nblocks; ++b) acc = acc + partial[4 * (size_t)b + k];
    out[k] = acc;
}

// Device address of `scratch` on the current device.
hipError_t scratch_address(double** out) {
    return hipGetSymbolAddress(reinterpret_cast<void**>(out), HIP_SYMBOL(scratch));
}

hipError_t launch_totals_sum(const double* partial, int nblocks, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_totals_sum, dim3(1), dim3(64), 0, s, partial, nblocks, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- launchers
// Grids: `block` threads per workgroup, a workgroup covers `block` cells
// (cell layout) or block / 4 cells (quad layout).
static inline int grid_cells(int n, int cells) { return (n + cells - 1) / cells; }

hipError_t launch_fill_u64(unsigned long long* p, int count, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3((count + 255) / 256), dim3(256), 0, s, p, count, v);
    return hipGetLastError();
}

hipError_t launch_dt(const DevMesh& M, const DevState& z, const Eos& eos, unsigned long long* slot, hipStream_t s) {
    hipLaunchKernelGGL(k_dt, dim3(grid_cells(M.n, kBlock)), dim3(kBlock), 0, s, M, z, eos, slot);
    return hipGetLastError();
}

hipError_t launch_flux1(int flux, int layout, int block, const DevMesh& M, const DevState& in, const DevState& out,
                        const Eos& eos, const StepCtl& ctl, hipStream_t s) {
    dim3 b(block);
    // HLLC_M, HLLC_LM, CIR1 and CIR2 run in the one-thread-per-cell layout only
    if (flux >= 4) layout = kLayoutCell;
    if (layout == kLayoutQuad) {
        dim3 g(grid_cells(M.n, block / 4));
        switch (flux) {
            case 0: hipLaunchKernelGGL(k_flux1_q<0>, g, b, 0, s, M, in, out, eos, ctl); break;
            case 1: hipLaunchKernelGGL(k_flux1_q<1>, g, b, 0, s, M, in, out, eos, ctl); break;
            case 2: hipLaunchKernelGGL(k_flux1_q<2>, g, b, 0, s, M, in, out, eos, ctl); break;
            default: hipLaunchKernelGGL(k_flux1_q<3>, g, b, 0, s, M, in, out, eos, ctl); break;
        }
        return hipGetLastError();
    }
    dim3 g(grid_cells(M.n, block));
    switch (flux) {
        case 0: hipLaunchKernelGGL(k_flux1<0>, g, b, 0, s, M, in, out, eos, ctl); break;
        case 1: hipLaunchKernelGGL(k_flux1<1>, g, b, 0, s, M, in, out, eos, ctl); break;
        case 2: hipLaunchKernelGGL(k_flux1<2>, g, b, 0, s, M, in, out, eos, ctl); break;
        case 4: hipLaunchKernelGGL(k_flux1<4>, g, b, 0, s, M, in, out, eos, ctl); break;
        case 5: hipLaunchKernelGGL(k_flux1<5>, g, b, 0, s, M, in, out, eos, ctl); break;
        case 6: hipLaunchKernelGGL(k_flux1<6>, g, b, 0, s, M, in, out, eos, ctl); break;
        case 7: hipLaunchKernelGGL(k_flux1<7>, g, b, 0, s, M, in, out, eos, ctl); break;
        default: hipLaunchKernelGGL(k_flux1<3>, g, b, 0, s, M, in, out, eos, ctl); break;
    }
    return hipGetLastError();
}

template <bool PRED, int LAYOUT>
static void launch_gp_t(int limiter, dim3 g, dim3 b, hipStream_t s, const DevMesh& M, const DevState& in,
                        const DevGrad& G, const DevState& half, const int* wait, const Eos& eos,
                        const StepCtl& ctl, double* raw) {
#define ZFV_GP(L)                                                                                             \
    do {                                                                                                      \
        if (LAYOUT == kLayoutQuad)                                                                            \
            hipLaunchKernelGGL((k_grad_pred_q<L, PRED>), g, b, 0, s, M, in, G, half, wait, eos, ctl, raw);    \
        else                                                                                                  \
            hipLaunchKernelGGL((k_grad_pred<L, PRED, LAYOUT == kLayoutBatch>), g, b, 0, s, M, in, G, half, wait, \
                               eos, ctl, raw);                                                                \
    } while (0)
    switch (limiter) {
        case 0: ZFV_GP(0); break;
        case 1: ZFV_GP(1); break;
        case 2: ZFV_GP(2); break;
        case 3: ZFV_GP(3); break;
        case 4: ZFV_GP(4); break;
        case 5: ZFV_GP(5); break;
        case 6: ZFV_GP(6); break;
        case 7: ZFV_GP(7); break;
        case 8: ZFV_GP(8); break;
        case 9: ZFV_GP(9); break;
        case 10: ZFV_GP(10); break;
        case 11: ZFV_GP(11); break;
        default: ZFV_GP(12); break;
    }
#undef ZFV_GP
}

hipError_t launch_grad_pred(int limiter, bool predictor, int layout, int block, const DevMesh& M, const DevState& in,
                            const DevGrad& G, const DevState& half, const int* wait, const Eos& eos,
                            const StepCtl& ctl, hipStream_t s, double* raw) {
    dim3 b(block);
    if (layout == kLayoutQuad) {
        dim3 g(grid_cells(M.n, block / 4));
        if (predictor) launch_gp_t<true, kLayoutQuad>(limiter, g, b, s, M, in, G, half, wait, eos, ctl, nullptr);
        else launch_gp_t<false, kLayoutQuad>(limiter, g, b, s, M, in, G, half, wait, eos, ctl, raw);
    } else if (layout == kLayoutBatch && predictor) {
        dim3 g(grid_cells(M.n, block));
        launch_gp_t<true, kLayoutBatch>(limiter, g, b, s, M, in, G, half, wait, eos, ctl, nullptr);
    } else {
        // without the predictor (the refinement criterion's gradient, on the
        // mesh's own order) the one-thread-per-cell kernel measures faster:
        // it keeps three waves per SIMD against the batched kernel's two
        dim3 g(grid_cells(M.n, block));
        if (predictor) launch_gp_t<true, kLayoutCell>(limiter, g, b, s, M, in, G, half, wait, eos, ctl, nullptr);
        else launch_gp_t<false, kLayoutCell>(limiter, g, b, s, M, in, G, half, wait, eos, ctl, raw);
    }
    return hipGetLastError();
}

hipError_t launch_corr(int flux, int layout, int block, const DevMesh& M, const DevState& init, const DevState& half,
                       const DevGrad& G, const int* wait_in, int* wait_out, const DevState& next,
                       const Eos& eos, const StepCtl& ctl, hipStream_t s) {
    dim3 b(block);
#define ZFV_CORR(K, ...) hipLaunchKernelGGL((K<__VA_ARGS__>), g, b, 0, s, M, init, half, G, wait_in, wait_out, next, eos, ctl)
    if (flux >= 4) layout = kLayoutCell;   // as in launch_flux1
    if (layout == kLayoutQuad) {
        dim3 g(grid_cells(M.n, block / 4));
        switch (flux) {
            case 0: ZFV_CORR(k_corr_q, 0); break;
            case 1: ZFV_CORR(k_corr_q, 1); break;
            case 2: ZFV_CORR(k_corr_q, 2); break;
            default: ZFV_CORR(k_corr_q, 3); break;
        }
    } else if (layout == kLayoutBatch) {
        dim3 g(grid_cells(M.n, block));
        switch (flux) {
            case 0: ZFV_CORR(k_corr, 0, true); break;
            case 1: ZFV_CORR(k_corr, 1, true); break;
            case 2: ZFV_CORR(k_corr, 2, true); break;
            default: ZFV_CORR(k_corr, 3, true); break;
        }
    } else {
        dim3 g(grid_cells(M.n, block));
        switch (flux) {
            case 0: ZFV_CORR(k_corr, 0, false); break;
            case 1: ZFV_CORR(k_corr, 1, false); break;
            case 2: ZFV_CORR(k_corr, 2, false); break;
            case 4: ZFV_CORR(k_corr, 4, false); break;
            case 5: ZFV_CORR(k_corr, 5, false); break;
            case 6: ZFV_CORR(k_corr, 6, false); break;
            case 7: ZFV_CORR(k_corr, 7, false); break;
            default: ZFV_CORR(k_corr, 3, false); break;
        }
    }
#undef ZFV_CORR
    return hipGetLastError();
}

hipError_t launch_totals(const DevMesh& M, const DevState& z, double* partial, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_totals, dim3(nblocks), dim3(kBlock), 0, s, M, z, partial);
    return hipGetLastError();
}

}  // namespace zfv
