// HIP kernels of the multi-material first-order step (MmFluid::update with
// acc 1, math/solver/mm_fluid.cpp:110-145, 166-194, 373-445): the PLIC
// interface reconstruction (:484-519), the composite Riemann fluxes of the
// faces between mixed cells (CrpMode::PLIC, the default) and the update.
//
// Same structure as the single-material step: one thread per cell gathering
// its faces in slot order, writing only itself; the next step's CFL
// candidate (compute_dt: min over the cell's faces of V / (S (|v.n| + c)))
// is fused into the update, reduced per workgroup and combined with one
// 64-bit atomicMin.  The mixture closure (a Newton iteration for mixed
// cells) runs once per cell per step for the new state and once for its
// sound speed, which is cached with the state: the reference recomputes the
// same value from the same state in every face flux.
#include "zfv_internal.hpp"
#include "zfv_step_common.hpp"
#include "zfv_mm.hpp"

namespace zfv {

using namespace mm;

namespace {

template <int NM>
__device__ __forceinline__ MZ<NM> load_mz(const MmLayer& L, int c, int st) {
    MZ<NM> z;
    z.r = L.r[c];
    z.u = L.u[c];
    z.v = L.v[c];
    z.P = L.P[c];
    z.e = L.e[c];
    z.T = L.T[c];
    z.c = L.c[c];
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        z.b[i] = L.b[i * st + c];
        z.rh[i] = L.rh[i * st + c];
    }
    return z;
}

template <int NM>
__device__ __forceinline__ void store_mz(const MmLayer& L, int c, int st, const MZ<NM>& z) {
    L.r[c] = z.r;
    L.u[c] = z.u;
    L.v[c] = z.v;
    L.P[c] = z.P;
    L.e[c] = z.e;
    L.T[c] = z.T;
    L.c[c] = z.c;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        L.b[i * st + c] = z.b[i];
        L.rh[i * st + c] = z.rh[i];
    }
}

// compute_dt candidate of one cell (mm_fluid.cpp:166-194).
template <int NM>
__device__ __forceinline__ double mm_dt_cand(const DevMesh& M, int c, const MZ<NM>& z) {
    double dt = 1.7976931348623157e308;   // numeric_limits<double>::max()
    const int code = M.code[c];
    const double V = M.vol[c];
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {
        if (((code >> (4 * s)) & 7) == SC_UNDEF) continue;
        const int so = s * M.stride + c;
        const double vn = z.u * M.nx[so] + z.v * M.ny[so];
        const double lambda = fabs(vn) + z.c;
        dt = smin(dt, V / (M.area[so] * lambda));
    }
    return dt;
}

template <int NM>
__device__ __forceinline__ void rotate_uv(const Frame& f, MZ<NM>& z) {
    const Z w = frame_local(f, Z{z.r, z.u, z.v, z.P, z.e});
    z.u = w.u;
    z.v = w.v;
}

}  // namespace

// Cells from {rho, u, v, P, beta} in place: the problem files'
// PState(rho, v, P, Fractions{...}, mixture) (models.cpp:176-185; Fractions
// normalise on construction, fractions.cpp:6-27), their sound speed and the
// first step's CFL candidate.
template <int NM>
__global__ void __launch_bounds__(kBlock) k_mm_prepare(DevMesh M, MmLayer L, Mix mx, unsigned long long* slot) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    double cand = INFINITY;
    if (c < tl.end) {
        const int st = M.stride;
        MZ<NM> z;
        z.r = L.r[c];
        z.u = L.u[c];
        z.v = L.v[c];
        z.P = L.P[c];
#pragma unroll
        for (int i = 0; i < NM; ++i) z.b[i] = L.b[i * st + c];
        frac_cutoff<NM>(z.b, kFracMinimal);
        get_reT<NM>(mx, z.r, z.P, z.b, NAN, nullptr, z.rh, z.e, z.T);
        z.c = sound_speed<NM>(mx, z.r, z.P, z.b, z.T, z.rh);
        store_mz<NM>(L, c, st, z);
        cand = mm_dt_cand<NM>(M, c, z);
    }
    block_min_atomic(slot + 3 * tl.pid, cand);
}

// Whether a face's flux is a composite Riemann flux: CRP on, and not both
// face states pure (MmFluid::calc_flux, mm_fluid.cpp:339-371).
template <int NM>
__device__ __forceinline__ bool crp_face(const double* CF, const MZ<NM>& zm, const MZ<NM>& zp) {
    return CF != nullptr && !(frac_index<NM>(zm.b) >= 0 && frac_index<NM>(zp.b) >= 0);
}

// Faces that need a composite Riemann flux (CrpMode::PLIC): a mixed cell on
// either side (MmFluid::calc_flux, mm_fluid.cpp:339-371; a boundary face's
// other side is the cell's own boundary value).  Appends slot-major pairs
// s * stride + c to `list`; `count` starts at zero.
template <int NM>
__global__ void __launch_bounds__(kBlock) k_mm_crp_mark(DevMesh M, MmLayer in, int* __restrict__ list,
                                                        unsigned int* count) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    if (c >= tl.end) return;
    const int st = M.stride;
    double b[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) b[i] = in.b[i * st + c];
    const bool own_pure = frac_index<NM>(b) >= 0;
    const int code = M.code[c];
    for (int s = 0; s < 8; ++s) {
        const int t = (code >> (4 * s)) & 7;
        if (t == SC_UNDEF) continue;
        bool nb_pure = own_pure;
        if (t <= SC_PERIODIC) {
            const int nb = M.adj[s * st + c];
            double bn[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i) bn[i] = in.b[i * st + nb];
            nb_pure = frac_index<NM>(bn) >= 0;
        }
        if (!(own_pure && nb_pure)) list[atomicAdd(count, 1u)] = s * st + c;
    }
}

// The composite Riemann flux of each listed face, one face per thread: the
// face-frame states, the master material
// (master_component, mm_fluid.cpp:201-226), the PLIC face fraction of the
// upwind cell's interface (alpha_sigma, :241-254, with the normals of
// interface_recovery) and calc_crp_flux (:259-337), rotated to the global
// frame, into CF[slot][k][stride].  These faces are few and need many more
// registers than the plain HLLC faces of k_mm_flux1, hence the compaction.
template <int NM>
__global__ void __launch_bounds__(64) k_mm_crp(DevMesh M, MmLayer in, Mix mx, StepCtl ctl,
                                            const double* __restrict__ inx, const double* __restrict__ iny,
                                            const int* __restrict__ list, const unsigned int* count,
                                            double* __restrict__ CF) {
    constexpr int K = 4 + NM;
    const int st = M.stride;
    const int n_faces = (int)*count;
    // L active lanes per one-wavefront workgroup: the faces' branchy paths
    // (closures, wave configurations) diverge, and a wavefront executes the
    // union of its lanes' paths, so few faces are spread over many
    // wavefronts; a long list fills whole wavefronts.
    int L = 4;
    while (L < 64 && (long long)L * gridDim.x < n_faces) L <<= 1;
    const int lane = (int)threadIdx.x;
    bool bad = false;
    for (int e = (int)blockIdx.x * L + lane; lane < L && e < n_faces; e += (int)gridDim.x * L) {
        const int so = list[e];
        const int s = so / st, c = so - s * st;
        const double dt = step_dt(ctl, 0);
        const MZ<NM> zc = load_mz<NM>(in, c, st);
        const int nib = (M.code[c] >> (4 * s)) & 15;
        const int t = nib & 7;
        const bool bound = t > SC_PERIODIC;
        const int nb = bound ? c : M.adj[so];
        const Frame fr = slot_frame(M, c, s, nib);
        MZ<NM> zp;
        if (!bound) {
            zp = load_mz<NM>(in, nb, st);
        } else {
            zp = zc;
            const Z w = frame_boundary(fr, Z{zc.r, zc.u, zc.v, zc.P, zc.e}, t == SC_WALL);
            zp.u = w.u;
            zp.v = w.v;
        }
        MZ<NM> zm = zc;
        rotate_uv<NM>(fr, zm);
        rotate_uv<NM>(fr, zp);
        const double S = M.area[so];
        const double h_L = M.vol[c] / S;
        const double h_R = (bound ? M.vol[c] : M.vol[nb]) / S;
        const int iA = master_component<NM>(zm, zp);
        const double a_L = alpha_of<NM>(zm, iA), a_R = alpha_of<NM>(zp, iA);
        const double vi = 0.5 * (zm.u + zp.u);
        const bool upwind = vi > 0.0;
        const double CFL = dt * fabs(vi) / (upwind ? h_L : h_R);
        double ncx = 0.0, ncy = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i)
            if (i == iA) {
                ncx = upwind ? inx[i * st + c] : -inx[i * st + nb];
                ncy = upwind ? iny[i * st + c] : -iny[i * st + nb];
            }
        const double cs = M.nx[so] * ncx + M.ny[so] * ncy;
        const double a_sig = f_between(average_flux(upwind ? a_L : a_R, cs, CFL), a_L, a_R);
        double F[K];
        calc_crp_flux<NM>(zm, zp, h_L, h_R, iA, a_sig, dt, mx, F, bad);
        const Q g = frame_global(fr, Q{F[0], F[1], F[2], F[3]});
        F[1] = g.mx;
        F[2] = g.my;
#pragma unroll
        for (int k = 0; k < K; ++k) CF[(s * K + k) * st + c] = F[k];
    }
    if (bad) atomicOr(ctl.err, 1);
}

// MmFluid::fluxes (mm_fluid.cpp:373-445): per face the mmf HLLC flux of the
// face-frame states (own, neighbour or boundary value, :98-108) or the
// composite flux from CF, q -= dt / V sum F S, the new state by the PT
// closure from the cell's previous P, T and true densities
// (models.cpp:187-201), its sound speed and the next step's CFL candidate.
template <int NM>
__global__ void __launch_bounds__(kBlock) k_mm_flux1(DevMesh M, MmLayer in, MmLayer out, Mix mx, StepCtl ctl,
                                                     const double* __restrict__ CF) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    const double dt = step_dt(ctl, tl.pid);
    step_prologue(ctl, tl, dt);
    double cand = INFINITY;
    if (c < tl.end) {
        const int st = M.stride;
        const MZ<NM> zc = load_mz<NM>(in, c, st);
        const int code = M.code[c];
        double acc[4 + NM];
#pragma unroll
        for (int k = 0; k < 4 + NM; ++k) acc[k] = 0.0;
        bool bad = false;
#pragma unroll 1
        for (int s = 0; s < 8; ++s) {
            const int nib = (code >> (4 * s)) & 15;
            const int t = nib & 7;
            if (t == SC_UNDEF) continue;
            const int so = s * st + c;
            const Frame fr = slot_frame(M, c, s, nib);
            MZ<NM> zp;
            if (t <= SC_PERIODIC) {
                zp = load_mz<NM>(in, M.adj[so], st);
            } else {
                zp = zc;
                const Z w = frame_boundary(fr, Z{zc.r, zc.u, zc.v, zc.P, zc.e}, t == SC_WALL);
                zp.u = w.u;
                zp.v = w.v;
            }
            MZ<NM> zm = zc;
            rotate_uv<NM>(fr, zm);
            rotate_uv<NM>(fr, zp);
            double F[4 + NM];
            if (crp_face<NM>(CF, zm, zp)) {
#pragma unroll
                for (int k = 0; k < 4 + NM; ++k) F[k] = CF[(s * (4 + NM) + k) * st + c];
            } else {
                hllc_mm<NM>(zm, zp, F, bad);
                const Q g = frame_global(fr, Q{F[0], F[1], F[2], F[3]});
                F[1] = g.mx;
                F[2] = g.my;
            }
            const double S = M.area[so];
#pragma unroll
            for (int k = 0; k < 4 + NM; ++k) acc[k] = acc[k] + F[k] * S;
        }
        double q[4 + NM];
        m_qstate<NM>(zc, q);
        const double coef = dt / M.vol[c];
#pragma unroll
        for (int k = 0; k < 4 + NM; ++k) q[k] = q[k] - coef * acc[k];
        if (q[0] < 0.0) bad = true;   // mm_fluid.cpp:440-442 throws
        MZ<NM> zn;
        zn.r = q[0];
        zn.u = q[1] / zn.r;
        zn.v = q[2] / zn.r;
        zn.e = q[3] / zn.r - 0.5 * (zn.u * zn.u + zn.v * zn.v);
#pragma unroll
        for (int i = 0; i < NM; ++i) zn.b[i] = q[4 + i] / zn.r;
        frac_cutoff<NM>(zn.b, 1.0e-12);
        frac_cutoff<NM>(zn.b, kFracMinimal);
        get_rPT<NM>(mx, zn.r, zn.e, zn.b, zc.P, zc.T, zc.rh, zn.rh, zn.P, zn.T);
        zn.c = sound_speed<NM>(mx, zn.r, zn.P, zn.b, zn.T, zn.rh);
        store_mz<NM>(out, c, st, zn);
        if (bad) atomicOr(ctl.err, 1);
        cand = mm_dt_cand<NM>(M, c, zn);
    }
    block_min_atomic(&ctl.dtmin[3 * tl.pid + ctl.wr], cand);
}

// MmFluid::interface_recovery (mm_fluid.cpp:484-519): per mixed cell and
// present material, n = -sum_faces face_fraction(alpha_c, alpha_n) S n_f,
// normalised, and the section position p = quad_find_section(alpha_c, n,
// hx, hy) (geom/sections.cpp:25-37); zero for pure cells.
template <int NM>
__global__ void __launch_bounds__(kBlock) k_mm_interface(DevMesh M, MmLayer L, const double* __restrict__ vx,
                                                         const double* __restrict__ vy, double* __restrict__ onx,
                                                         double* __restrict__ ony, double* __restrict__ op) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= M.n) return;
    const int st = M.stride;
    const MZ<NM> z = load_mz<NM>(L, c, st);
    double nsx[NM], nsy[NM], ps[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        nsx[i] = 0.0;
        nsy[i] = 0.0;
        ps[i] = 0.0;
    }
    if (!frac_pure<NM>(z.b)) {
        double a_c[NM];
#pragma unroll
        for (int i = 0; i < NM; ++i) a_c[i] = z.b[i] > 0.0 ? z.b[i] * z.r / z.rh[i] : 0.0;   // volume_fractions
        const int code = M.code[c];
#pragma unroll 1
        for (int s = 0; s < 8; ++s) {
            const int t = (code >> (4 * s)) & 7;
            if (t == SC_UNDEF) continue;
            const int so = s * st + c;
            const double a = M.area[so];
            const double Sx = a * M.nx[so], Sy = a * M.ny[so];
            const int nb = t >= SC_WALL ? c : M.adj[so];   // neib() is the cell itself at a boundary
            const double rn = L.r[nb];
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                if (!(a_c[i] > 0.0)) continue;
                const double rhn = L.rh[i * st + nb];
                double a_n = 0.0;   // PState::alpha (models.cpp:224-227)
                if (!isnan(rhn)) {
                    const double x = L.b[i * st + nb] * rn / rhn;
                    const double lo = (1.0 < x) ? 1.0 : x;
                    a_n = (0.0 < lo) ? lo : 0.0;
                }
                const double ff = face_fraction(a_c[i], a_n);
                nsx[i] = nsx[i] - ff * Sx;
                nsy[i] = nsy[i] - ff * Sy;
            }
        }
        const double hx = sqrt((vx[5 * st + c] - vx[3 * st + c]) * (vx[5 * st + c] - vx[3 * st + c]) +
                               (vy[5 * st + c] - vy[3 * st + c]) * (vy[5 * st + c] - vy[3 * st + c]));
        const double hy = sqrt((vx[7 * st + c] - vx[1 * st + c]) * (vx[7 * st + c] - vx[1 * st + c]) +
                               (vy[7 * st + c] - vy[1 * st + c]) * (vy[7 * st + c] - vy[1 * st + c]));
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            if (!(a_c[i] > 0.0)) continue;
            const double zz = nsx[i] * nsx[i] + nsy[i] * nsy[i];
            if (zz > 0.0) {
                const double r = sqrt(zz);
                nsx[i] = nsx[i] / r;
                nsy[i] = nsy[i] / r;
            }
            ps[i] = quad_find_section(a_c[i], nsx[i], nsy[i], hx, hy);
        }
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        onx[i * st + c] = nsx[i];
        ony[i * st + c] = nsy[i];
        op[i * st + c] = ps[i];
    }
}

// ---- launchers --------------------------------------------------------------
static inline dim3 mm_grid(int n) { return dim3((n + kBlock - 1) / kBlock); }

#define ZFV_MM_DISPATCH(NMAT, KERNEL, ...)                                                              \
    switch (NMAT) {                                                                                     \
        case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                                       \
        case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                                       \
        case 3: hipLaunchKernelGGL(KERNEL<3>, __VA_ARGS__); break;                                       \
        default: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                                      \
    }

hipError_t mm_prepare(const DevMesh& M, const MmLayer& L, const Mix& mx, unsigned long long* slot, hipStream_t s) {
    ZFV_MM_DISPATCH(mx.n, k_mm_prepare, mm_grid(M.n), dim3(kBlock), 0, s, M, L, mx, slot);
    return hipGetLastError();
}

hipError_t mm_flux1(const DevMesh& M, const MmLayer& in, const MmLayer& out, const Mix& mx, const StepCtl& ctl,
                    const double* CF, hipStream_t s) {
    ZFV_MM_DISPATCH(mx.n, k_mm_flux1, mm_grid(M.n), dim3(kBlock), 0, s, M, in, out, mx, ctl, CF);
    return hipGetLastError();
}

__global__ void __launch_bounds__(64) k_zero_u32(unsigned int* p) {
    if (threadIdx.x == 0) *p = 0u;
}

hipError_t mm_crp_mark(const DevMesh& M, const MmLayer& in, const Mix& mx, int* list, unsigned int* count,
                       hipStream_t s) {
    // the counter is cleared by a kernel, not a memset node, so the step can be captured
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, count);
    ZFV_MM_DISPATCH(mx.n, k_mm_crp_mark, mm_grid(M.n), dim3(kBlock), 0, s, M, in, list, count);
    return hipGetLastError();
}

hipError_t mm_crp(const DevMesh& M, const MmLayer& in, const Mix& mx, const StepCtl& ctl, const double* nx,
                  const double* ny, const int* list, const unsigned int* count, double* CF, hipStream_t s) {
    // 2048 one-wavefront workgroups, grid-stride over any list length
    ZFV_MM_DISPATCH(mx.n, k_mm_crp, dim3(2048), dim3(64), 0, s, M, in, mx, ctl, nx, ny, list, count, CF);
    return hipGetLastError();
}

hipError_t mm_interface(const DevMesh& M, const MmLayer& L, const Mix& mx, const double* vx, const double* vy,
                        double* nx, double* ny, double* p, hipStream_t s) {
    ZFV_MM_DISPATCH(mx.n, k_mm_interface, mm_grid(M.n), dim3(kBlock), 0, s, M, L, vx, vy, nx, ny, p);
    return hipGetLastError();
}

}  // namespace zfv
