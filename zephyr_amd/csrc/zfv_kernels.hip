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

namespace zfv {

namespace {

constexpr unsigned long long kInfBits = 0x7ff0000000000000ULL;   // +inf

__device__ __forceinline__ int xcd_block(int b, int G) {
    const int x = b & 7, q = G >> 3, r = G & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ __forceinline__ int cell_of_thread() {
    return xcd_block((int)blockIdx.x, (int)gridDim.x) * kBlock + (int)threadIdx.x;
}

__device__ __forceinline__ Z load_z(const DevState& s, int c) {
    return Z{s.r[c], s.u[c], s.v[c], s.p[c], s.e[c]};
}

__device__ __forceinline__ void store_z(const DevState& s, int c, const Z& z) {
    s.r[c] = z.r;
    s.u[c] = z.u;
    s.v[c] = z.v;
    s.p[c] = z.p;
    s.e[c] = z.e;
}

// h / (|v| + c) (sm_fluid.cpp:105-110, incircle precomputed per cell).
__device__ __forceinline__ double cell_dt(const Z& z, double h, const Eos& eos) {
    double c = eos_c(eos, z.r, z.p);
    return h / (sqrt(vsq(z)) + c);
}

__device__ __forceinline__ double step_dt(const StepCtl& ctl) {
    double dmin = __longlong_as_double((long long)ctl.dtmin[ctl.rd]);
    return smin(ctl.cfl * dmin, ctl.max_dt);
}

__device__ __forceinline__ void step_prologue(const StepCtl& ctl, double dt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *ctl.dt_out = dt;
        ctl.dtmin[ctl.clr] = kInfBits;
    }
}

__device__ __forceinline__ void block_min_atomic(unsigned long long* slot, double v) {
    if (!(v < INFINITY)) v = INFINITY;   // NaN and +inf never win
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    __shared__ double red[kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
#pragma unroll
        for (int i = 1; i < kBlock / 64; ++i) m = (red[i] < m) ? red[i] : m;
        if (m < INFINITY) atomicMin(slot, (unsigned long long)__double_as_longlong(m));
    }
}

// Frame of slot s of cell c (code nibble `nib`).
__device__ __forceinline__ Frame slot_frame(const DevMesh& M, int c, int s, int nib) {
    const bool gen = (nib & kGeneralNormal) != 0;
    double nx = 0.0, ny = 0.0;
    if (gen) {
        nx = M.nx[s * M.stride + c];
        ny = M.ny[s * M.stride + c];
    }
    return frame_of(s & 3, gen, nx, ny);
}

// Neighbour state through a slot and the LSM offset dr (gradient.h:245-252).
__device__ __forceinline__ void lsm_neib(const DevMesh& M, const DevState& in, int c, int s,
                                         int t, const Frame& f, const Z& zc, double ccx, double ccy,
                                         Z& zn, double& drx, double& dry) {
    const int nb = M.adj[s * M.stride + c];
    if (t == SC_INNER || t == SC_PERIODIC) {
        zn.r = in.r[nb];
        zn.u = in.u[nb];
        zn.v = in.v[nb];
        zn.p = in.p[nb];
        zn.e = 0.0;
    } else {
        zn = frame_boundary(f, zc, t == SC_WALL);
    }
    double px, py;
    if (t == SC_INNER) {
        px = M.cx[nb];
        py = M.cy[nb];
    } else {
        frame_symm(f, ccx, ccy, M.fcx[s * M.stride + c], M.fcy[s * M.stride + c], px, py);
    }
    drx = px - ccx;
    dry = py - ccy;
}

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

}  // namespace

__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long* p, int count, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_dt(DevMesh M, DevState z, Eos eos, unsigned long long* slot) {
    const int c = cell_of_thread();
    double cand = INFINITY;
    if (c < M.n) cand = cell_dt(load_z(z, c), M.h[c], eos);
    block_min_atomic(slot, cand);
}

// SmFluid::fluxes, first order (sm_fluid.cpp:127-174), plus the next step's
// CFL candidate.
template <int FLUX>
__global__ void __launch_bounds__(kBlock) k_flux1(DevMesh M, DevState in, DevState out, Eos eos, StepCtl ctl) {
    const int n = M.n;
    const int c = cell_of_thread();
    const double dt = step_dt(ctl);
    step_prologue(ctl, dt);
    double cand = INFINITY;
    if (c < n) {
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
        const Z zn = to_z(q, eos);
        store_z(out, c, zn);
        if (bad) *ctl.err = 1;
        cand = cell_dt(zn, M.h[c], eos);
    }
    block_min_atomic(&ctl.dtmin[ctl.wr], cand);
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
// gradient of rho ([2][stride]) for the refinement criterion.
template <int LIM, bool PRED>
__global__ void __launch_bounds__(kBlock) k_grad_pred(DevMesh M, DevState in, DevGrad G, DevState half,
                                                      const int* __restrict__ wait, Eos eos, StepCtl ctl,
                                                      double* __restrict__ raw) {
    const int n = M.n;
    const int c = cell_of_thread();
    const double dt = step_dt(ctl);
    if (PRED) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl.dtmin[ctl.clr] = kInfBits;
    }
    if (c >= n) return;

    const Z zc = load_z(in, c);
    const double ccx = M.cx[c], ccy = M.cy[c];
    const int code = M.code[c];

    double A00 = 0.0, A01 = 0.0, A10 = 0.0, A11 = 0.0;
    double Fx[4] = {0.0, 0.0, 0.0, 0.0}, Fy[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < 8; ++s) {
        const int nib = (code >> (4 * s)) & 15;
        const int t = nib & 7;
        if (t == SC_UNDEF) continue;
        const Frame fr = slot_frame(M, c, s, nib);
        Z zn;
        double drx, dry, dnx, dny;
        lsm_neib(M, in, c, s, t, fr, zc, ccx, ccy, zn, drx, dry);
        lsm_drn(fr, drx, dry, dnx, dny);
        const double w = M.w[s * M.stride + c];
        const double dF[4] = {zn.r - zc.r, zn.u - zc.u, zn.v - zc.v, zn.p - zc.p};
        if (fr.gen || fr.side < 2) {
            const double wd = w * dnx;
            A00 = A00 + wd * drx;
            A01 = A01 + wd * dry;
#pragma unroll
            for (int k = 0; k < 4; ++k) Fx[k] = Fx[k] + (w * dF[k]) * dnx;
        }
        if (fr.gen || fr.side >= 2) {
            const double wd = w * dny;
            A10 = A10 + wd * drx;
            A11 = A11 + wd * dry;
#pragma unroll
            for (int k = 0; k < 4; ++k) Fy[k] = Fy[k] + (w * dF[k]) * dny;
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
    for (int s = 0; s < 8; ++s) {
        const int nib = (code >> (4 * s)) & 15;
        const int t = nib & 7;
        if (t == SC_UNDEF) continue;
        const Frame fr = slot_frame(M, c, s, nib);
        Z zn;
        double drx, dry, dnx, dny;
        lsm_neib(M, in, c, s, t, fr, zc, ccx, ccy, zn, drx, dry);
        lsm_drn(fr, drx, dry, dnx, dny);
        const double w = M.w[s * M.stride + c];
        const double dFp[4] = {zn.r - zc.r, zn.u - zc.u, zn.v - zc.v, zn.p - zc.p};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double du_n = gx[k] * drx + gy[k] * dry;
            const double dFm = 2.0 * du_n - dFp[k];
            const double lim = limiter2<LIM>(dFm, dFp[k]) * dFp[k];
            if (fr.gen || fr.side < 2) Lx[k] = Lx[k] + (w * lim) * dnx;
            if (fr.gen || fr.side >= 2) Ly[k] = Ly[k] + (w * lim) * dny;
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
    if (wait[c] <= 0) {
        double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
        for (int s = 0; s < 8; ++s) {
            const int nib = (code >> (4 * s)) & 15;
            if ((nib & 7) == SC_UNDEF) continue;
            const Frame fr = slot_frame(M, c, s, nib);
            const double dx = M.fcx[s * M.stride + c] - ccx;
            const double dy = M.fcy[s * M.stride + c] - ccy;
            Z zm = extrap(zc, gx, gy, dx, dy);
            zm.e = eos_e(eos, zm.r, zm.p);
            if (is_bad(zm, eos)) zm = zc;
            Q f = phys_flux(frame_local(fr, zm));
            f = frame_global(fr, f);
            const double a = M.area[s * M.stride + c];
            F0 = F0 + f.r * a;
            F1 = F1 + f.mx * a;
            F2 = F2 + f.my * a;
            F3 = F3 + f.E * a;
        }
        Q q = to_q(zc);
        const double coef = 0.5 * dt / M.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        zh = to_z(q, eos);
        if (is_bad(zh, eos)) zh = zc;
    }
    store_z(half, c, zh);
}

// SmFluid::fluxes_stage2, corrector (sm_fluid.cpp:255-366), plus the next
// step's CFL candidate.  `wait` is double-buffered: neighbours' isolation is
// read from the pre-pass array.
template <int FLUX>
__global__ void __launch_bounds__(kBlock) k_corr(DevMesh M, DevState init, DevState half, DevGrad G,
                                                 const int* __restrict__ wait_in, int* __restrict__ wait_out,
                                                 DevState next, Eos eos, StepCtl ctl) {
    const int n = M.n;
    const int c = cell_of_thread();
    const double dt = step_dt(ctl);
    if (blockIdx.x == 0 && threadIdx.x == 0) *ctl.dt_out = dt;
    double cand = INFINITY;
    if (c < n) {
        const Z zc = load_z(init, c);
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
        for (int s = 0; s < 8; ++s) {
            const int nib = (code >> (4 * s)) & 15;
            const int t = nib & 7;
            if (t == SC_UNDEF) continue;
            const Frame fr = slot_frame(M, c, s, nib);
            const int nb = M.adj[s * M.stride + c];
            const bool bound = t >= SC_WALL;
            const bool wall = t == SC_WALL;
            const Z znh = bound ? frame_boundary(fr, zch, wall) : load_z(half, nb);
            const bool robust = wc > 0 || wait_in[nb] > 0;
            Z zm, zp;
            if (robust) {
                zm = zch;
                zp = znh;
            } else {
                const double fx = M.fcx[s * M.stride + c], fy = M.fcy[s * M.stride + c];
                zm = extrap(zch, gx, gy, fx - ccx, fy - ccy);
                zm.e = eos_e(eos, zm.r, zm.p);
                if (is_bad(zm, eos)) zm = zch;
                if (!bound) {
                    double nx, ny;
                    if (t == SC_INNER) {
                        nx = M.cx[nb];
                        ny = M.cy[nb];
                    } else {
                        frame_symm(fr, ccx, ccy, fx, fy, nx, ny);
                    }
                    double ngx[4], ngy[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        ngx[k] = G.dx[k * M.stride + nb];
                        ngy[k] = G.dy[k * M.stride + nb];
                    }
                    zp = extrap(znh, ngx, ngy, fx - nx, fy - ny);
                    zp.e = eos_e(eos, zp.r, zp.p);
                    if (is_bad(zp, eos)) zp = znh;
                } else {
                    zp = frame_boundary(fr, zm, wall);
                }
            }
            const Z lm = frame_local(fr, zm), lp = frame_local(fr, zp);
            Q f = robust ? flux_hll(lm, lp, eos, bad) : num_flux<FLUX>(lm, lp, eos, bad);
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
        Z zn = to_z(q, eos);
        int wn;
        if (is_bad(zn, eos)) {
            zn = zc;
            wn = 3;
        } else {
            wn = wc > 0 ? wc - 1 : wc;
        }
        store_z(next, c, zn);
        wait_out[c] = wn;
        if (bad) *ctl.err = 1;
        cand = cell_dt(zn, M.h[c], eos);
    }
    block_min_atomic(&ctl.dtmin[ctl.wr], cand);
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

// ---------------------------------------------------------------- launchers
static inline int grid_of(int n) { return (n + kBlock - 1) / kBlock; }

hipError_t launch_fill_u64(unsigned long long* p, int count, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3((count + 255) / 256), dim3(256), 0, s, p, count, v);
    return hipGetLastError();
}

hipError_t launch_dt(const DevMesh& M, const DevState& z, const Eos& eos, unsigned long long* slot, hipStream_t s) {
    hipLaunchKernelGGL(k_dt, dim3(grid_of(M.n)), dim3(kBlock), 0, s, M, z, eos, slot);
    return hipGetLastError();
}

hipError_t launch_flux1(int flux, const DevMesh& M, const DevState& in, const DevState& out,
                        const Eos& eos, const StepCtl& ctl, hipStream_t s) {
    dim3 g(grid_of(M.n)), b(kBlock);
    switch (flux) {
        case 0: hipLaunchKernelGGL(k_flux1<0>, g, b, 0, s, M, in, out, eos, ctl); break;
        case 1: hipLaunchKernelGGL(k_flux1<1>, g, b, 0, s, M, in, out, eos, ctl); break;
        case 2: hipLaunchKernelGGL(k_flux1<2>, g, b, 0, s, M, in, out, eos, ctl); break;
        default: hipLaunchKernelGGL(k_flux1<3>, g, b, 0, s, M, in, out, eos, ctl); break;
    }
    return hipGetLastError();
}

template <bool PRED>
static void launch_gp_t(int limiter, dim3 g, dim3 b, hipStream_t s, const DevMesh& M, const DevState& in,
                        const DevGrad& G, const DevState& half, const int* wait, const Eos& eos,
                        const StepCtl& ctl, double* raw) {
#define ZFV_GP(L) hipLaunchKernelGGL((k_grad_pred<L, PRED>), g, b, 0, s, M, in, G, half, wait, eos, ctl, raw)
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

hipError_t launch_grad_pred(int limiter, bool predictor, const DevMesh& M, const DevState& in, const DevGrad& G,
                            const DevState& half, const int* wait, const Eos& eos, const StepCtl& ctl,
                            hipStream_t s, double* raw) {
    dim3 g(grid_of(M.n)), b(kBlock);
    if (predictor) launch_gp_t<true>(limiter, g, b, s, M, in, G, half, wait, eos, ctl, nullptr);
    else launch_gp_t<false>(limiter, g, b, s, M, in, G, half, wait, eos, ctl, raw);
    return hipGetLastError();
}

hipError_t launch_corr(int flux, const DevMesh& M, const DevState& init, const DevState& half,
                       const DevGrad& G, const int* wait_in, int* wait_out, const DevState& next,
                       const Eos& eos, const StepCtl& ctl, hipStream_t s) {
    dim3 g(grid_of(M.n)), b(kBlock);
    switch (flux) {
        case 0: hipLaunchKernelGGL(k_corr<0>, g, b, 0, s, M, init, half, G, wait_in, wait_out, next, eos, ctl); break;
        case 1: hipLaunchKernelGGL(k_corr<1>, g, b, 0, s, M, init, half, G, wait_in, wait_out, next, eos, ctl); break;
        case 2: hipLaunchKernelGGL(k_corr<2>, g, b, 0, s, M, init, half, G, wait_in, wait_out, next, eos, ctl); break;
        default: hipLaunchKernelGGL(k_corr<3>, g, b, 0, s, M, init, half, G, wait_in, wait_out, next, eos, ctl); break;
    }
    return hipGetLastError();
}

hipError_t launch_totals(const DevMesh& M, const DevState& z, double* partial, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_totals, dim3(nblocks), dim3(kBlock), 0, s, M, z, partial);
    return hipGetLastError();
}

}  // namespace zfv
