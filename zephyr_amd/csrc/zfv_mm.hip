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

// EuCell::faces(dir) (amr_faces.cpp:186-197): whether slot so is swept.
__device__ __forceinline__ bool mm_swept(const DevMesh& M, int so, int dir) {
    if (dir == 1) return !(fabs(M.nx[so]) < 0.7);
    if (dir == 2) return !(fabs(M.ny[so]) < 0.7);
    return true;
}

// The new state from the conserved vector q: PState(q, mixture, P0, T0,
// rhos0) (models.cpp:187-201) with the cell's previous P, T and true
// densities as the closure's starting point.  The sound speed is left to
// the caller.
template <int NM>
__device__ __forceinline__ MZ<NM> mm_from_q(const Mix& mx, const double* q, const MZ<NM>& z0) {
    MZ<NM> zn;
    zn.r = q[0];
    zn.u = q[1] / zn.r;
    zn.v = q[2] / zn.r;
    zn.e = q[3] / zn.r - 0.5 * (zn.u * zn.u + zn.v * zn.v);
#pragma unroll
    for (int i = 0; i < NM; ++i) zn.b[i] = q[4 + i] / zn.r;
    frac_cutoff<NM>(zn.b, 1.0e-12);
    frac_cutoff<NM>(zn.b, kFracMinimal);
    get_rPT<NM>(mx, zn.r, zn.e, zn.b, z0.P, z0.T, z0.rh, zn.rh, zn.P, zn.T);
    zn.c = 0.0;
    return zn;
}

// ---- second order (acc 2) ----------------------------------------------------
// Gradient rows of k_mm_grad: rho, u, v, e, beta[NM], P, T, rhos[NM].
template <int NM>
__device__ __forceinline__ double mg_du(const double* gx, const double* gy, int k, int c, int st, double dx,
                                        double dy) {
    return gx[k * st + c] * dx + gy[k * st + c] * dy;
}

// FaceExtra::Direct m() / p() (face_extra.h:122-184) of the half state of
// cell c: every component moved by its gradient along (dx, dy), then
// PState::interpolation_update (models.cpp:381-391): mass fractions
// normalised, e, T and the true densities re-closed from (rho, P) starting
// from the extrapolated T and densities.  The extrapolated e is overwritten
// there, so it is not formed.  (The reference's g_z dr.z term is a signed
// zero here: dr.z = 0.)  With `sound`, also the sound speed the fluxes use
// (sound_speed_rP from the re-closed T and densities).
template <int NM>
__device__ __forceinline__ MZ<NM> mm_face_state(const MmLayer& H, const Mix& mx, const double* gx, const double* gy,
                                                int c, int st, double dx, double dy, bool sound) {
    MZ<NM> z;
    z.r = H.r[c] + mg_du<NM>(gx, gy, 0, c, st, dx, dy);
    z.u = H.u[c] + mg_du<NM>(gx, gy, 1, c, st, dx, dy);
    z.v = H.v[c] + mg_du<NM>(gx, gy, 2, c, st, dx, dy);
    z.P = H.P[c] + mg_du<NM>(gx, gy, 4 + NM, c, st, dx, dy);
    const double T0 = H.T[c] + mg_du<NM>(gx, gy, 5 + NM, c, st, dx, dy);
    double rh0[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        z.b[i] = H.b[i * st + c] + mg_du<NM>(gx, gy, 4 + i, c, st, dx, dy);
        rh0[i] = H.rh[i * st + c] + mg_du<NM>(gx, gy, 6 + NM + i, c, st, dx, dy);
    }
    frac_cutoff<NM>(z.b, kFracMinimal);
    get_reT<NM>(mx, z.r, z.P, z.b, T0, rh0, z.rh, z.e, z.T);
    z.c = sound ? sound_speed<NM>(mx, z.r, z.P, z.b, z.T, z.rh) : 0.0;
    return z;
}

// Material index of the face state of mm_face_state (only its mass
// fractions: the same sums and normalisation).
template <int NM>
__device__ __forceinline__ int mm_face_index(const MmLayer& H, const double* gx, const double* gy, int c, int st,
                                             double dx, double dy) {
    double b[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) b[i] = H.b[i * st + c] + mg_du<NM>(gx, gy, 4 + i, c, st, dx, dy);
    frac_cutoff<NM>(b, kFracMinimal);
    return frac_index<NM>(b);
}

// The reference's boundary_value (mm_fluid.cpp:98-108) of a face state:
// the state with the velocity reflected at walls.
template <int NM>
__device__ __forceinline__ MZ<NM> mm_boundary(const Frame& fr, const MZ<NM>& z, bool wall) {
    MZ<NM> r = z;
    const Z w = frame_boundary(fr, Z{z.r, z.u, z.v, z.P, z.e}, wall);
    r.u = w.u;
    r.v = w.v;
    return r;
}

// The two states of slot s of cell c in the global frame: acc 1 the cells'
// own (MmFluid::fluxes, :373-445), acc 2 the extrapolated half states of
// fluxes_stage2 (:580-680; a periodic face's neighbour centre is the
// neighbour's own, since face.is_boundary() is false there).
template <int NM, int ACC>
__device__ __forceinline__ void mm_face_pair(const DevMesh& M, const MmLayer& in, const MmLayer& half, const Mix& mx,
                                             const double* gx, const double* gy, int c, int so, int t,
                                             const Frame& fr, MZ<NM>& zm, MZ<NM>& zp) {
    const int st = M.stride;
    const bool inner = t <= SC_PERIODIC;
    const int nb = inner ? M.adj[so] : c;
    if (ACC == 1) {
        zm = load_mz<NM>(in, c, st);
        zp = inner ? load_mz<NM>(in, nb, st) : mm_boundary<NM>(fr, zm, t == SC_WALL);
    } else {
        const double fx = M.fcx[so], fy = M.fcy[so];
        zm = mm_face_state<NM>(half, mx, gx, gy, c, st, fx - M.cx[c], fy - M.cy[c], true);
        zp = inner ? mm_face_state<NM>(half, mx, gx, gy, nb, st, fx - M.cx[nb], fy - M.cy[nb], true)
                   : mm_boundary<NM>(fr, zm, t == SC_WALL);
    }
}

// Whether slot s of cell c is a composite-flux face: not both face states
// pure (MmFluid::calc_flux, :347-371).
template <int NM, int ACC>
__device__ __forceinline__ bool mm_crp_slot(const DevMesh& M, const MmLayer& in, const MmLayer& half,
                                            const double* gx, const double* gy, int c, int so, int t) {
    const int st = M.stride;
    const bool inner = t <= SC_PERIODIC;
    const int nb = inner ? M.adj[so] : c;
    if (ACC == 1) {
        double b[NM];
#pragma unroll
        for (int i = 0; i < NM; ++i) b[i] = in.b[i * st + c];
        if (frac_index<NM>(b) < 0) return true;
        if (!inner) return false;
#pragma unroll
        for (int i = 0; i < NM; ++i) b[i] = in.b[i * st + nb];
        return frac_index<NM>(b) < 0;
    }
    const double fx = M.fcx[so], fy = M.fcy[so];
    if (mm_face_index<NM>(half, gx, gy, c, st, fx - M.cx[c], fy - M.cy[c]) < 0) return true;
    return inner && mm_face_index<NM>(half, gx, gy, nb, st, fx - M.cx[nb], fy - M.cy[nb]) < 0;
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

// Faces that need a composite Riemann flux (CrpMode::PLIC): a mixed face
// state on either side (MmFluid::calc_flux, mm_fluid.cpp:347-371; a
// boundary face's other side is the boundary value of the cell's own).  acc
// 1: the cells' states; acc 2: the extrapolated, normalised half states.
// Appends slot-major pairs s * stride + c to `list` (counter count[0]) and,
// when `cells` is given, each cell with such a face to `cells` (counter
// count[1]); the counters start at zero.
template <int NM, int ACC>
__global__ void __launch_bounds__(kBlock) k_mm_crp_mark(DevMesh M, MmLayer in, MmLayer half,
                                                        const double* __restrict__ gx, const double* __restrict__ gy,
                                                        int dir, int* __restrict__ list, int* __restrict__ cells,
                                                        unsigned int* count) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    if (c >= tl.end) return;
    const int st = M.stride;
    const int code = M.code[c];
    bool any = false;
    for (int s = 0; s < 8; ++s) {
        const int t = (code >> (4 * s)) & 7;
        if (t == SC_UNDEF || !mm_swept(M, s * st + c, dir)) continue;
        if (mm_crp_slot<NM, ACC>(M, in, half, gx, gy, c, s * st + c, t)) {
            list[atomicAdd(count, 1u)] = s * st + c;
            any = true;
        }
    }
    if (any && cells) cells[atomicAdd(count + 1, 1u)] = c;
}

// The composite Riemann flux of each listed face, one face per thread: the
// face-frame states, the master material
// (master_component, mm_fluid.cpp:201-226), the PLIC face fraction of the
// upwind cell's interface (alpha_sigma, :241-254, with the normals of
// interface_recovery) and calc_crp_flux (:259-337), rotated to the global
// frame, into CF[slot][k][stride].  These faces are few and need many more
// registers than the plain HLLC faces of k_mm_flux1, hence the compaction.
template <int NM, int ACC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) k_mm_crp(DevMesh M, MmLayer in, MmLayer half, const double* __restrict__ gx,
                                            const double* __restrict__ gy, Mix mx, StepCtl ctl, MmOpts opt,
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
        const double dt = opt.dt_scale * step_dt(ctl, 0);
        const int nib = (M.code[c] >> (4 * s)) & 15;
        const int t = nib & 7;
        const bool bound = t > SC_PERIODIC;
        const int nb = bound ? c : M.adj[so];
        const Frame fr = slot_frame(M, c, s, nib);
        MZ<NM> zm, zp;
        mm_face_pair<NM, ACC>(M, in, half, mx, gx, gy, c, so, t, fr, zm, zp);
        rotate_uv<NM>(fr, zm);
        rotate_uv<NM>(fr, zp);
        const double S = M.area[so];
        const double h_L = M.vol[c] / S;
        const double h_R = (bound ? M.vol[c] : M.vol[nb]) / S;
        const int iA = master_component<NM>(zm, zp);
        const double a_L = alpha_of<NM>(zm, iA), a_R = alpha_of<NM>(zp, iA);
        double a_sig;
        if (opt.crp == 4) {   // PLIC
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
            a_sig = f_between(average_flux(upwind ? a_L : a_R, cs, CFL), a_L, a_R);
        } else {   // the heuristic templates V3 / V5 / VS
            a_sig = opt.crp == 1 ? face_fraction_v3(a_L, a_R)
                  : opt.crp == 2 ? face_fraction_v5(a_L, a_R) : face_fraction_s(a_L, a_R);
        }
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

// MmFluid::fluxes (mm_fluid.cpp:373-445) of cell c: per face the mmf HLLC
// flux of the face-frame states (own, neighbour or boundary value, :98-108)
// or the composite flux from CF, q -= dt / V sum F S, the new state by the
// PT closure from the cell's previous P, T and true densities
// (models.cpp:187-201) and its sound speed; returns the next step's CFL
// candidate.  With `defer` a cell with a composite face is left alone
// (+inf, nothing stored, CF not read): k_mm_flux1_list updates it once CF
// is complete.  PLAIN: 1 = the plain faces take HLLC (the default, compiled
// alone), 2 = the flux of opt.flux.
template <int NM, int PLAIN>
__device__ __forceinline__ double mm_flux1_cell(const DevMesh& M, const MmLayer& in, const MmLayer& out,
                                                const Mix& mx, const StepCtl& ctl, const MmOpts& opt,
                                                const double* __restrict__ CF, int c, double dt, bool defer) {
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
        if (t == SC_UNDEF || !mm_swept(M, s * st + c, opt.dir)) continue;
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
            if (defer) return INFINITY;
#pragma unroll
            for (int k = 0; k < 4 + NM; ++k) F[k] = CF[(s * (4 + NM) + k) * st + c];
        } else {
            if (PLAIN == 1) hllc_mm<NM>(zm, zp, F, bad);
            else plain_flux_mm<NM>(opt.flux, zm, zp, F, bad);
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
    MZ<NM> zn = mm_from_q<NM>(mx, q, zc);
    zn.c = sound_speed<NM>(mx, zn.r, zn.P, zn.b, zn.T, zn.rh);
    store_mz<NM>(out, c, st, zn);
    if (bad) atomicOr(ctl.err, 1);
    return mm_dt_cand<NM>(M, c, zn);
}

// Every cell (defer = false), or every cell without a composite face
// (defer = true, while k_mm_crp fills CF on another stream).
template <int NM, int PLAIN>
__global__ void __launch_bounds__(kBlock) k_mm_flux1(DevMesh M, MmLayer in, MmLayer out, Mix mx, StepCtl ctl,
                                                     MmOpts opt, const double* __restrict__ CF, int defer) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    const double step = step_dt(ctl, tl.pid);
    step_prologue(ctl, tl, step);
    const double dt = opt.dt_scale * step;
    double cand = INFINITY;
    if (c < tl.end) cand = mm_flux1_cell<NM, PLAIN>(M, in, out, mx, ctl, opt, CF, c, dt, defer != 0);
    block_min_atomic(&ctl.dtmin[3 * tl.pid + ctl.wr], cand);
}

// The cells k_mm_flux1 deferred: cells[0 .. count[1]) from k_mm_crp_mark,
// grid-stride.  The multi-material solver runs one problem (dt slot 0).
template <int NM, int PLAIN>
__global__ void __launch_bounds__(kBlock) k_mm_flux1_list(DevMesh M, MmLayer in, MmLayer out, Mix mx, StepCtl ctl,
                                                          MmOpts opt, const double* __restrict__ CF,
                                                          const int* __restrict__ cells,
                                                          const unsigned int* __restrict__ count) {
    const double dt = opt.dt_scale * step_dt(ctl, 0);
    const int n = (int)count[1];
    double cand = INFINITY;
    for (int e = (int)(blockIdx.x * blockDim.x + threadIdx.x); e < n; e += (int)(gridDim.x * blockDim.x)) {
        const double v = mm_flux1_cell<NM, PLAIN>(M, in, out, mx, ctl, opt, CF, cells[e], dt, false);
        cand = v < cand ? v : cand;   // a NaN candidate never wins, as in block_min_atomic
    }
    block_min_atomic(&ctl.dtmin[ctl.wr], cand);
}

// MmFluid::fluxes_stage1 (mm_fluid.cpp:521-578): a mixed cell keeps its
// state; a pure cell advances half a step, q - 0.5 dt / V sum_faces
// Flux(z_m) S with the physical flux of its own extrapolated, re-closed
// face states, and re-closes from its previous P, T and true densities.
template <int NM>
__global__ void __launch_bounds__(kBlock) k_mm_stage1(DevMesh M, MmLayer in, MmLayer half, Mix mx, StepCtl ctl,
                                                      MmOpts opt, const double* __restrict__ gx,
                                                      const double* __restrict__ gy) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    if (c >= tl.end) return;
    const int st = M.stride;
    const MZ<NM> zc = load_mz<NM>(in, c, st);
    if (frac_index<NM>(zc.b) < 0) {
        store_mz<NM>(half, c, st, zc);
        return;
    }
    const double dt = opt.dt_scale * step_dt(ctl, tl.pid);
    const double ccx = M.cx[c], ccy = M.cy[c];
    const int code = M.code[c];
    double acc[4 + NM];
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) acc[k] = 0.0;
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {
        const int nib = (code >> (4 * s)) & 15;
        if ((nib & 7) == SC_UNDEF || !mm_swept(M, s * st + c, opt.dir)) continue;
        const int so = s * st + c;
        const Frame fr = slot_frame(M, c, s, nib);
        MZ<NM> zm = mm_face_state<NM>(in, mx, gx, gy, c, st, M.fcx[so] - ccx, M.fcy[so] - ccy, false);
        rotate_uv<NM>(fr, zm);
        double F[4 + NM];
        m_flux<NM>(zm, F);
        const Q g = frame_global(fr, Q{F[0], F[1], F[2], F[3]});
        F[1] = g.mx;
        F[2] = g.my;
        const double S = M.area[so];
#pragma unroll
        for (int k = 0; k < 4 + NM; ++k) acc[k] = acc[k] + F[k] * S;
    }
    double q[4 + NM];
    m_qstate<NM>(zc, q);
    const double coef = 0.5 * dt / M.vol[c];
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) q[k] = q[k] - coef * acc[k];
    store_mz<NM>(half, c, st, mm_from_q<NM>(mx, q, zc));
}

// MmFluid::fluxes_stage2 (mm_fluid.cpp:580-680) of cell c: per face the
// states extrapolated from the half layer with the gradients of the old one
// and re-closed, the mmf HLLC flux or the composite flux from CF, q -= dt /
// V sum F S on the old conserved vector, the closure from the old P, T and
// true densities and the sound speed; returns the next step's CFL
// candidate.  `defer` as in mm_flux1_cell.
template <int NM, int PLAIN>
__device__ __forceinline__ double mm_flux2_cell(const DevMesh& M, const MmLayer& in, const MmLayer& half,
                                                const MmLayer& out, const Mix& mx, const StepCtl& ctl,
                                                const MmOpts& opt, const double* __restrict__ gx,
                                                const double* __restrict__ gy, const double* __restrict__ CF,
                                                int c, double dt, bool defer) {
    const int st = M.stride;
    const int code = M.code[c];
    double acc[4 + NM];
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) acc[k] = 0.0;
    bool bad = false;
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {
        const int nib = (code >> (4 * s)) & 15;
        const int t = nib & 7;
        if (t == SC_UNDEF || !mm_swept(M, s * st + c, opt.dir)) continue;
        const int so = s * st + c;
        double F[4 + NM];
        if (CF != nullptr && mm_crp_slot<NM, 2>(M, in, half, gx, gy, c, so, t)) {
            if (defer) return INFINITY;
#pragma unroll
            for (int k = 0; k < 4 + NM; ++k) F[k] = CF[(s * (4 + NM) + k) * st + c];
        } else {
            const Frame fr = slot_frame(M, c, s, nib);
            MZ<NM> zm, zp;
            mm_face_pair<NM, 2>(M, in, half, mx, gx, gy, c, so, t, fr, zm, zp);
            rotate_uv<NM>(fr, zm);
            rotate_uv<NM>(fr, zp);
            if (PLAIN == 1) hllc_mm<NM>(zm, zp, F, bad);
            else plain_flux_mm<NM>(opt.flux, zm, zp, F, bad);
            const Q g = frame_global(fr, Q{F[0], F[1], F[2], F[3]});
            F[1] = g.mx;
            F[2] = g.my;
        }
        const double S = M.area[so];
#pragma unroll
        for (int k = 0; k < 4 + NM; ++k) acc[k] = acc[k] + F[k] * S;
    }
    const MZ<NM> zc = load_mz<NM>(in, c, st);
    double q[4 + NM];
    m_qstate<NM>(zc, q);
    const double coef = dt / M.vol[c];
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) q[k] = q[k] - coef * acc[k];
    if (q[0] < 0.0) bad = true;   // mm_fluid.cpp:671-673 throws
    MZ<NM> zn = mm_from_q<NM>(mx, q, zc);
    zn.c = sound_speed<NM>(mx, zn.r, zn.P, zn.b, zn.T, zn.rh);
    store_mz<NM>(out, c, st, zn);
    if (bad) atomicOr(ctl.err, 1);
    return mm_dt_cand<NM>(M, c, zn);
}

template <int NM, int PLAIN>
__global__ void __launch_bounds__(kBlock) k_mm_flux2(DevMesh M, MmLayer in, MmLayer half, MmLayer out, Mix mx,
                                                     StepCtl ctl, MmOpts opt, const double* __restrict__ gx,
                                                     const double* __restrict__ gy, const double* __restrict__ CF,
                                                     int defer) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    const double step = step_dt(ctl, tl.pid);
    step_prologue(ctl, tl, step);
    const double dt = opt.dt_scale * step;
    double cand = INFINITY;
    if (c < tl.end) cand = mm_flux2_cell<NM, PLAIN>(M, in, half, out, mx, ctl, opt, gx, gy, CF, c, dt, defer != 0);
    block_min_atomic(&ctl.dtmin[3 * tl.pid + ctl.wr], cand);
}

template <int NM, int PLAIN>
__global__ void __launch_bounds__(kBlock) k_mm_flux2_list(DevMesh M, MmLayer in, MmLayer half, MmLayer out, Mix mx,
                                                          StepCtl ctl, MmOpts opt, const double* __restrict__ gx,
                                                          const double* __restrict__ gy,
                                                          const double* __restrict__ CF,
                                                          const int* __restrict__ cells,
                                                          const unsigned int* __restrict__ count) {
    const double dt = opt.dt_scale * step_dt(ctl, 0);
    const int n = (int)count[1];
    double cand = INFINITY;
    for (int e = (int)(blockIdx.x * blockDim.x + threadIdx.x); e < n; e += (int)(gridDim.x * blockDim.x)) {
        const double v = mm_flux2_cell<NM, PLAIN>(M, in, half, out, mx, ctl, opt, gx, gy, CF, cells[e], dt, false);
        cand = v < cand ? v : cand;
    }
    block_min_atomic(&ctl.dtmin[ctl.wr], cand);
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

// ---- adaptation: MmFluid::compute_grad, set_flags, distributor --------------

// The limiters by number (the gradient runs once per adaptation).
__device__ double limiter_rt(int kind, double a, double b) {
    switch (kind) {
        case 0: return limiter2<0>(a, b);
        case 1: return limiter2<1>(a, b);
        case 2: return limiter2<2>(a, b);
        case 3: return limiter2<3>(a, b);
        case 4: return limiter2<4>(a, b);
        case 5: return limiter2<5>(a, b);
        case 6: return limiter2<6>(a, b);
        case 7: return limiter2<7>(a, b);
        case 8: return limiter2<8>(a, b);
        case 9: return limiter2<9>(a, b);
        case 10: return limiter2<10>(a, b);
        case 11: return limiter2<11>(a, b);
        default: return limiter2<12>(a, b);
    }
}

// Components of the gradient: rho, u, v, e, beta[i] (what the split
// reads), then P, T, rhos[i] (the second-order face states carry every
// PState component, face_extra.h:122-184).  A true density is NaN where its
// material is absent, and so is its gradient wherever the stencil meets
// such a cell.  Row k of a layer (k is uniform across the wavefront).
template <int NM>
__device__ __forceinline__ const double* mg_row(const MmLayer& L, int st, int k) {
    if (k < 4) return k == 0 ? L.r : k == 1 ? L.u : k == 2 ? L.v : L.e;
    if (k < 4 + NM) return L.b + (size_t)(k - 4) * st;
    if (k == 4 + NM) return L.P;
    if (k == 5 + NM) return L.T;
    return L.rh + (size_t)(k - 6 - NM) * st;
}

// MmFluid::compute_grad (mm_fluid.cpp:439-466): with composite Riemann
// problems on (crp != 0) zero in a mixed cell and in a cell with a mixed
// face neighbour; else gradient::LSM (gradient.h:226-282) and
// gradient::limiting (:412-474) with the mm boundary_value
// (mm_fluid.cpp:98-108: walls reflect the velocity), component by
// component as k_grad_pred does for SmFluid (the 2 x 2 block of the LSM
// system).  The first `rows` of the 6 + 2 NM components (all of them for a
// second-order step, the 4 + NM the split distributor reads for
// set_flags) go in chunks of KC through the two passes (the unlimited and
// the limited right-hand side), so a thread holds KC components' sums
// rather than all of them; the LSM matrix is built in the first chunk's
// first pass.  Each component's sums run over the slots in the same order
// as unchunked, so the result does not depend on KC.  gx, gy:
// [6 + 2 NM][stride].
template <int NM, int KC>
__global__ void __launch_bounds__(kBlock) k_mm_grad(DevMesh M, MmLayer L, int crp, int limiter, int rows,
                                                    double* __restrict__ gx, double* __restrict__ gy) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= M.n) return;
    const int st = M.stride;
    const int code = M.code[c];
    bool zero = false;
    if (crp) {
        double b[NM];
#pragma unroll
        for (int i = 0; i < NM; ++i) b[i] = L.b[i * st + c];
        zero = frac_index<NM>(b) < 0;
        for (int s = 0; s < 8 && !zero; ++s) {
            const int t = (code >> (4 * s)) & 7;
            if (t != SC_INNER && t != SC_PERIODIC) continue;
            const int nb = M.adj[s * st + c];
#pragma unroll
            for (int i = 0; i < NM; ++i) b[i] = L.b[i * st + nb];
            if (frac_index<NM>(b) < 0) zero = true;
        }
    }
    if (zero) {
        for (int k = 0; k < rows; ++k) {
            gx[k * st + c] = 0.0;
            gy[k * st + c] = 0.0;
        }
        return;
    }
    const double ccx = M.cx[c], ccy = M.cy[c];
    double A00 = 0.0, A01 = 0.0, A10 = 0.0, A11 = 0.0;
    double B00 = 0.0, B01 = 0.0, B10 = 0.0, B11 = 0.0;
#pragma unroll 1
    for (int k0 = 0; k0 < rows; k0 += KC) {
        double zc[KC], Gx[KC], Gy[KC];
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            zc[j] = k0 + j < rows ? mg_row<NM>(L, st, k0 + j)[c] : 0.0;
            Gx[j] = 0.0;
            Gy[j] = 0.0;
        }
        // pass 0: the unlimited right-hand side; pass 1: the limited one
#pragma unroll 1
        for (int pass = 0; pass < 2; ++pass) {
            double Fx[KC], Fy[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                Fx[j] = 0.0;
                Fy[j] = 0.0;
            }
#pragma unroll 1
            for (int s = 0; s < 8; ++s) {
                const int nib = (code >> (4 * s)) & 15;
                const int t = nib & 7;
                if (t == SC_UNDEF) continue;
                const int so = s * st + c;
                const Frame fr = slot_frame(M, c, s, nib);
                const int nb = M.adj[so];
                const bool inner = t == SC_INNER || t == SC_PERIODIC;
                double zn[KC];
#pragma unroll
                for (int j = 0; j < KC; ++j) zn[j] = (inner && k0 + j < rows) ? mg_row<NM>(L, st, k0 + j)[nb] : zc[j];
                if (!inner && k0 == 0) {   // the wall's reflected velocity (rows 1, 2 lie in the first chunk)
                    const Z w = frame_boundary(fr, Z{zc[0], zc[1], zc[2], 0.0, zc[3]}, t == SC_WALL);
                    zn[1] = w.u;
                    zn[2] = w.v;
                }
                double px = M.cx[nb], py = M.cy[nb];
                if (t != SC_INNER) frame_symm(fr, ccx, ccy, M.fcx[so], M.fcy[so], px, py);
                const double drx = px - ccx, dry = py - ccy;
                double dnx, dny;
                if (!fr.gen) {
                    dnx = fr.side < 2 ? drx : 0.0;
                    dny = fr.side < 2 ? 0.0 : dry;
                } else {
                    const double d = drx * fr.nx + dry * fr.ny;
                    dnx = d * fr.nx;
                    dny = d * fr.ny;
                }
                const bool xp = fr.gen || fr.side < 2, yp = fr.gen || fr.side >= 2;
                const double w = M.w[so];
                if (k0 == 0 && pass == 0) {   // the LSM matrix
                    if (xp) {
                        const double wd = w * dnx;
                        A00 = A00 + wd * drx;
                        A01 = A01 + wd * dry;
                    }
                    if (yp) {
                        const double wd = w * dny;
                        A10 = A10 + wd * drx;
                        A11 = A11 + wd * dry;
                    }
                }
#pragma unroll
                for (int j = 0; j < KC; ++j) {
                    const double dF = zn[j] - zc[j];
                    double term = dF;
                    if (pass == 1) {
                        const double du_n = Gx[j] * drx + Gy[j] * dry;
                        term = limiter_rt(limiter, 2.0 * du_n - dF, dF) * dF;
                    }
                    if (xp) Fx[j] = Fx[j] + (w * term) * dnx;
                    if (yp) Fy[j] = Fy[j] + (w * term) * dny;
                }
            }
            if (k0 == 0 && pass == 0) {   // its inverse
                if (fabs(A11) < 1e-8) A11 = 1.0;
                const double det = A11 * A00 + (-A01) * A10;
                const double invdet = 1.0 / det;
                B00 = A11 * invdet;
                B01 = (-A01) * invdet;
                B10 = (-A10) * invdet;
                B11 = A00 * invdet;
            }
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                Gx[j] = B00 * Fx[j] + B01 * Fy[j];
                Gy[j] = B10 * Fx[j] + B11 * Fy[j];
            }
        }
#pragma unroll
        for (int j = 0; j < KC; ++j)
            if (k0 + j < rows) {
                gx[(k0 + j) * st + c] = Gx[j];
                gy[(k0 + j) * st + c] = Gy[j];
            }
    }
}

// MmFluid::set_flags (mm_fluid.cpp:836-885) after its compute_grad: +1 for
// a mixed cell or a pure cell with a face neighbour of another material
// (or mixed; boundary faces skipped), else -1.
template <int NM>
__global__ void __launch_bounds__(kBlock) k_mm_flags(DevMesh M, MmLayer L, int* __restrict__ flag) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= M.n) return;
    const int st = M.stride;
    double b[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) b[i] = L.b[i * st + c];
    if (!frac_pure<NM>(b)) {
        flag[c] = 1;
        return;
    }
    const int idx = frac_index<NM>(b);
    const int code = M.code[c];
    bool split = false;
    for (int s = 0; s < 8 && !split; ++s) {
        const int t = (code >> (4 * s)) & 7;
        if (t != SC_INNER && t != SC_PERIODIC) continue;
        const int nb = M.adj[s * st + c];
#pragma unroll
        for (int i = 0; i < NM; ++i) b[i] = L.b[i * st + nb];
        if (frac_index<NM>(b) != idx) split = true;
    }
    flag[c] = split ? 1 : -1;
}

// MmFluid::distributor split (mm_fluid.cpp:768-800), one thread per child
// (ref_list: the refined cells; child i of parent ip at next[ip] + i, its
// geometry already written): child = parent + grad . dr with momentum- and
// energy-consistent corrections, beta normalised, P from pressure_re(rho,
// e, beta, {P0, T0} of the parent), T from temperature_rP(rho, P, beta,
// {T0}); the true densities are the default-constructed ScalarSet, zeros.
// The sound speed is left to k_mm_cache.  A state the reference rejects
// (it throws) sets bit 0 of the error word.
template <int NM>
__global__ void __launch_bounds__(256) k_mm_split(AmrMesh A, MmLayer L, const double* __restrict__ gx,
                                                  const double* __restrict__ gy, Mix mx,
                                                  const int* __restrict__ ref_list, int n_refine, int* err) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int pi = t >> 2, i = t & 3;
    if (pi >= n_refine) return;
    const int st = A.stride;
    const int ip = ref_list[pi];
    const int c = A.next[ip] + i;
    const MZ<NM> zp = load_mz<NM>(L, ip, st);
    const double drx = A.cx[c] - A.cx[ip], dry = A.cy[c] - A.cy[ip];
    double d[4 + NM];
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) d[k] = gx[k * st + ip] * drx + gy[k * st + ip] * dry;
    MZ<NM> z;
    z.r = zp.r + d[0];
    const double ratio = zp.r / z.r;
    z.u = zp.u + ratio * d[1];
    z.v = zp.v + ratio * d[2];
#pragma unroll
    for (int k = 0; k < NM; ++k) z.b[k] = zp.b[k] + (zp.r / z.r) * d[4 + k];
    frac_cutoff<NM>(z.b, kFracMinimal);
    const double du = zp.u - z.u, dv = zp.v - z.v;
    z.e = (zp.e + 0.5 * (du * du + dv * dv)) + (zp.r / z.r) * d[3];
    const int idx = frac_index<NM>(z.b);
    double rh[NM];
    if (idx >= 0) {
        const Mat m = pick<NM>(mx, idx);
        z.P = sg_pressure_re(m, z.r, z.e);
        z.T = sg_temperature_rP(m, z.r, z.P);
    } else {
        double Tt;
        DPDT vl, el;
        find_rPT<NM>(mx, z.r, z.e, z.b, zp.P, zp.T, nullptr, rh, z.P, Tt, vl, el);
        DPDT v;
        find_rT_rP<NM>(mx, z.r, z.P, z.b, zp.T, nullptr, rh, z.T, v);
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) z.rh[k] = 0.0;
    z.c = 0.0;
    if (mm_is_bad<NM>(z)) atomicOr(err, 1);
    store_mz<NM>(L, c, st, z);
}

// MmFluid::distributor merge (mm_fluid.cpp:802-830), one thread per
// coarsened family (fam: its four children, parent at next[child]):
// volume-weighted conserved sum, negative partial densities clipped, the
// closure started from the volume-weighted mean P and T (reference
// densities as the true-density guesses).
template <int NM>
__global__ void __launch_bounds__(256) k_mm_merge(AmrMesh A, MmLayer L, Mix mx, const int* __restrict__ fam,
                                                  int n_parents, int* err) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= n_parents) return;
    const int st = A.stride;
    double q[4 + NM];
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) q[k] = 0.0;
    double mean_p = 0.0, mean_t = 0.0;
    const int ip = A.next[fam[4 * t]];
    for (int k = 0; k < 4; ++k) {
        const int c = fam[4 * t + k];
        const MZ<NM> zc = load_mz<NM>(L, c, st);
        double qc[4 + NM];
        m_qstate<NM>(zc, qc);
        const double V = A.vol[c];
#pragma unroll
        for (int j = 0; j < 4 + NM; ++j) q[j] = q[j] + qc[j] * V;
        mean_p = mean_p + zc.P * V;
        mean_t = mean_t + zc.T * V;
    }
    const double Vp = A.vol[ip];
#pragma unroll
    for (int j = 0; j < 4 + NM; ++j) q[j] = q[j] / Vp;
    mean_p = mean_p / Vp;
    mean_t = mean_t / Vp;
#pragma unroll
    for (int k = 0; k < NM; ++k) q[4 + k] = (0.0 < q[4 + k]) ? q[4 + k] : 0.0;
    MZ<NM> z;   // PState(QState, mixture, P0, T0) (models.cpp:187-201)
    z.r = q[0];
    z.u = q[1] / z.r;
    z.v = q[2] / z.r;
    z.e = q[3] / z.r - 0.5 * (z.u * z.u + z.v * z.v);
#pragma unroll
    for (int k = 0; k < NM; ++k) z.b[k] = q[4 + k] / z.r;
    frac_cutoff<NM>(z.b, 1.0e-12);
    frac_cutoff<NM>(z.b, kFracMinimal);
    get_rPT<NM>(mx, z.r, z.e, z.b, mean_p, mean_t, nullptr, z.rh, z.P, z.T);
    z.c = 0.0;
    if (mm_is_bad<NM>(z)) atomicOr(err, 1);
    store_mz<NM>(L, ip, st, z);
}

// SwapLists::move_elements (swap_lists.h:129-138) for the state rows.
template <int NM>
__global__ void __launch_bounds__(256) k_mm_move(MmLayer L, int st, int n_short, int nz_total, int n,
                                                 const AmrMarks* __restrict__ rank, int bound,
                                                 const int* __restrict__ swap_src, const int* __restrict__ swap_dst) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= bound) return;
    const int n_swap = n_short >= n ? nz_total : rank[n_short].nz;   // the holes
    if (i >= n_swap) return;
    const int from = swap_src[i], to = swap_dst[i];
    store_mz<NM>(L, to, st, load_mz<NM>(L, from, st));
}

// The cached sound speed (sound_speed_rP with the state's T and true
// densities as guesses, as compute_dt and the fluxes take it) and the
// compute_dt candidate of every cell after an adaptation.
template <int NM>
__global__ void __launch_bounds__(kBlock) k_mm_cache(DevMesh M, MmLayer L, Mix mx, unsigned long long* slot) {
    const Tile tl = tile_of(M, (int)blockDim.x);
    const int c = tl.c0 + (int)threadIdx.x;
    double cand = INFINITY;
    if (c < tl.end) {
        const int st = M.stride;
        MZ<NM> z = load_mz<NM>(L, c, st);
        z.c = sound_speed<NM>(mx, z.r, z.P, z.b, z.T, z.rh);
        L.c[c] = z.c;
        cand = mm_dt_cand<NM>(M, c, z);
    }
    block_min_atomic(slot + 3 * tl.pid, cand);
}

// ---- launchers --------------------------------------------------------------
static inline dim3 mm_grid(int n) { return dim3((n + kBlock - 1) / kBlock); }
// workgroups of the deferred-cell passes (a few thousand cells at most, grid-stride)
constexpr int kListGroups = 64;

#define ZFV_MM_DISPATCH(NMAT, KERNEL, ...)                                                              \
    switch (NMAT) {                                                                                     \
        case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                                       \
        case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                                       \
        case 3: hipLaunchKernelGGL(KERNEL<3>, __VA_ARGS__); break;                                       \
        default: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                                      \
    }

// KERNEL<NM, V>, V = 1 or 2 (the order of accuracy, or the plain-face flux
// variant of the update kernels).
#define ZFV_MM_DISPATCH_ACC(NMAT, ACC, KERNEL, ...)                                                     \
    switch ((NMAT) * 2 + (ACC) - 1) {                                                                   \
        case 2: hipLaunchKernelGGL((KERNEL<1, 1>), __VA_ARGS__); break;                                  \
        case 3: hipLaunchKernelGGL((KERNEL<1, 2>), __VA_ARGS__); break;                                  \
        case 4: hipLaunchKernelGGL((KERNEL<2, 1>), __VA_ARGS__); break;                                  \
        case 5: hipLaunchKernelGGL((KERNEL<2, 2>), __VA_ARGS__); break;                                  \
        case 6: hipLaunchKernelGGL((KERNEL<3, 1>), __VA_ARGS__); break;                                  \
        case 7: hipLaunchKernelGGL((KERNEL<3, 2>), __VA_ARGS__); break;                                  \
        case 8: hipLaunchKernelGGL((KERNEL<4, 1>), __VA_ARGS__); break;                                  \
        default: hipLaunchKernelGGL((KERNEL<4, 2>), __VA_ARGS__); break;                                 \
    }

hipError_t mm_prepare(const DevMesh& M, const MmLayer& L, const Mix& mx, unsigned long long* slot, hipStream_t s) {
    ZFV_MM_DISPATCH(mx.n, k_mm_prepare, mm_grid(M.n), dim3(kBlock), 0, s, M, L, mx, slot);
    return hipGetLastError();
}

hipError_t mm_flux1(const DevMesh& M, const MmLayer& in, const MmLayer& out, const Mix& mx, const StepCtl& ctl,
                    const MmOpts& o, const double* CF, int pass, const int* cells, const unsigned int* count,
                    hipStream_t s) {
    const int plain = o.flux == kFluxHllc ? 1 : 2;
    if (pass == 2) {
        ZFV_MM_DISPATCH_ACC(mx.n, plain, k_mm_flux1_list, dim3(kListGroups), dim3(kBlock), 0, s, M, in, out, mx, ctl,
                            o, CF, cells, count);
    } else {
        ZFV_MM_DISPATCH_ACC(mx.n, plain, k_mm_flux1, mm_grid(M.n), dim3(kBlock), 0, s, M, in, out, mx, ctl, o, CF,
                            pass);
    }
    return hipGetLastError();
}

__global__ void __launch_bounds__(64) k_zero_u32(unsigned int* p) {
    if (threadIdx.x < 2) p[threadIdx.x] = 0u;
}

hipError_t mm_crp_mark(const DevMesh& M, const MmLayer& in, const MmLayer& half, const Mix& mx, int acc,
                       const MmOpts& o, const double* gx, const double* gy, int* list, int* cells,
                       unsigned int* count, hipStream_t s) {
    // the counters are cleared by a kernel, not a memset node, so the step can be captured
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, count);
    ZFV_MM_DISPATCH_ACC(mx.n, acc, k_mm_crp_mark, mm_grid(M.n), dim3(kBlock), 0, s, M, in, half, gx, gy, o.dir,
                        list, cells, count);
    return hipGetLastError();
}

hipError_t mm_crp(const DevMesh& M, const MmLayer& in, const MmLayer& half, const Mix& mx, int acc,
                  const MmOpts& o, const double* gx, const double* gy, const StepCtl& ctl, const double* nx,
                  const double* ny, const int* list, const unsigned int* count, double* CF, int groups,
                  hipStream_t s) {
    // `groups` one-wavefront workgroups, grid-stride over any list length
    ZFV_MM_DISPATCH_ACC(mx.n, acc, k_mm_crp, dim3(groups), dim3(64), 0, s, M, in, half, gx, gy, mx, ctl, o, nx, ny,
                        list, count, CF);
    return hipGetLastError();
}

hipError_t mm_stage1(const DevMesh& M, const MmLayer& in, const MmLayer& half, const Mix& mx, const StepCtl& ctl,
                     const MmOpts& o, const double* gx, const double* gy, hipStream_t s) {
    ZFV_MM_DISPATCH(mx.n, k_mm_stage1, mm_grid(M.n), dim3(kBlock), 0, s, M, in, half, mx, ctl, o, gx, gy);
    return hipGetLastError();
}

hipError_t mm_flux2(const DevMesh& M, const MmLayer& in, const MmLayer& half, const MmLayer& out, const Mix& mx,
                    const StepCtl& ctl, const MmOpts& o, const double* gx, const double* gy, const double* CF,
                    int pass, const int* cells, const unsigned int* count, hipStream_t s) {
    const int plain = o.flux == kFluxHllc ? 1 : 2;
    if (pass == 2) {
        ZFV_MM_DISPATCH_ACC(mx.n, plain, k_mm_flux2_list, dim3(kListGroups), dim3(kBlock), 0, s, M, in, half, out,
                            mx, ctl, o, gx, gy, CF, cells, count);
    } else {
        ZFV_MM_DISPATCH_ACC(mx.n, plain, k_mm_flux2, mm_grid(M.n), dim3(kBlock), 0, s, M, in, half, out, mx, ctl, o,
                            gx, gy, CF, pass);
    }
    return hipGetLastError();
}

hipError_t mm_interface(const DevMesh& M, const MmLayer& L, const Mix& mx, const double* vx, const double* vy,
                        double* nx, double* ny, double* p, hipStream_t s) {
    ZFV_MM_DISPATCH(mx.n, k_mm_interface, mm_grid(M.n), dim3(kBlock), 0, s, M, L, vx, vy, nx, ny, p);
    return hipGetLastError();
}


hipError_t mm_grad(const DevMesh& M, const MmLayer& L, const Mix& mx, int crp, int limiter, int rows, double* gx,
                   double* gy, hipStream_t s) {
    // the rows of set_flags (4 + NM) in one chunk; all 6 + 2 NM of a
    // second-order step in chunks of 4, which keeps more wavefronts resident
    const bool split = rows <= 4 + mx.n;
#define ZFV_MM_GRAD(NMC)                                                                                     \
    hipLaunchKernelGGL((split ? k_mm_grad<NMC, 4 + NMC> : k_mm_grad<NMC, 4>), mm_grid(M.n), dim3(kBlock), 0, s, \
                       M, L, crp, limiter, rows, gx, gy)
    switch (mx.n) {
        case 1: ZFV_MM_GRAD(1); break;
        case 2: ZFV_MM_GRAD(2); break;
        case 3: ZFV_MM_GRAD(3); break;
        default: ZFV_MM_GRAD(4); break;
    }
#undef ZFV_MM_GRAD
    return hipGetLastError();
}

hipError_t mm_flags(const DevMesh& M, const MmLayer& L, const Mix& mx, int* flag, hipStream_t s) {
    ZFV_MM_DISPATCH(mx.n, k_mm_flags, mm_grid(M.n), dim3(kBlock), 0, s, M, L, flag);
    return hipGetLastError();
}

hipError_t mm_distribute(const AmrMesh& A, const MmLayer& L, const Mix& mx, const double* gx, const double* gy,
                         const int* ref_list, int n_refine, const int* fam, int n_parents, int* err, hipStream_t s) {
    if (n_refine > 0)
        ZFV_MM_DISPATCH(mx.n, k_mm_split, dim3((4 * n_refine + 255) / 256), dim3(256), 0, s, A, L, gx, gy, mx,
                        ref_list, n_refine, err);
    if (n_parents > 0)
        ZFV_MM_DISPATCH(mx.n, k_mm_merge, dim3((n_parents + 255) / 256), dim3(256), 0, s, A, L, mx, fam, n_parents,
                        err);
    return hipGetLastError();
}

hipError_t mm_move(const MmLayer& L, const Mix& mx, int stride, int n, int n_short, int nz_total, const AmrMarks* rank,
                   int bound, const int* swap_src, const int* swap_dst, hipStream_t s) {
    if (bound <= 0) return hipSuccess;
    ZFV_MM_DISPATCH(mx.n, k_mm_move, dim3((bound + 255) / 256), dim3(256), 0, s, L, stride, n_short, nz_total, n,
                    rank, bound, swap_src, swap_dst);
    return hipGetLastError();
}

hipError_t mm_cache(const DevMesh& M, const MmLayer& L, const Mix& mx, unsigned long long* slot, hipStream_t s) {
    ZFV_MM_DISPATCH(mx.n, k_mm_cache, mm_grid(M.n), dim3(kBlock), 0, s, M, L, mx, slot);
    return hipGetLastError();
}

}  // namespace zfv
