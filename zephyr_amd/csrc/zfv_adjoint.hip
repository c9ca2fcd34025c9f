// Discrete adjoint of the MUSCL-Hancock + HLLC step (SmFluid::update,
// acc 2) on Cartesian meshes: the exact reverse-mode derivative of the
// device step's operations, for the gradient of a final-time functional
// with respect to the initial fields (no reference counterpart).
//
// One adjoint step maps lambda' = dJ/dz^{n+1} to lambda = dJ/dz^n through
// the step's parts in reverse:
//   corrector   z' = PState(Q(z) - dt/V sum F(zm(h_c, G_c), zp(h_n, G_n)) S)
//   predictor   h  = PState(Q(z) - dt/(2V) sum Flux(zm(z, G)) S)
//   gradient    G  = LSM + limiting of {rho, u, v, p} over the 1-ring
//   time step   dt = min(CFL min_c h_c / (|v_c| + c_c), max_dt)
// Branches (supersonic HLLC exits, bad-state fallbacks, limiter regions,
// min/max selections) take the derivative of the branch the forward step
// took.  Each corrector face is reversed once, by the cell that owns it,
// which writes the adjoints of both face states into per-slot buffers (on a
// Cartesian mesh slot s of cell c points to n and slot s ^ 1 of n points
// back to c); the gradient's neighbour contributions are written per slot
// and gathered through the opposite slot.  No floating-point atomics are
// needed except for the scalar dt adjoint.
#include "zfv_internal.hpp"
#include "zfv_adjoint_ops.hpp"

namespace zfv {

using namespace adj;

namespace {

__device__ __forceinline__ Z load_zs(const DevState& s, int c) { return Z{s.r[c], s.u[c], s.v[c], s.p[c], s.e[c]}; }

__device__ __forceinline__ int cell_id_a() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }

// The step's dt as the forward kernels compute it (step_dt): min(CFL * dmin, max_dt).
struct DtRef {
    const unsigned long long* slot;   // min_c h / (|v| + c) as a positive double's bits
    double cfl;
    const double* max_dt;
};
__device__ __forceinline__ double dt_of(const DtRef& d) {
    return smin(d.cfl * __longlong_as_double((long long)*d.slot), *d.max_dt);
}

__device__ __forceinline__ void block_sum_atomic(double* out, double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double red[8];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
        if (s != 0.0) atomicAdd(out, s);
    }
}

}  // namespace

// ---- adjoint kernels --------------------------------------------------------
// Buffers: W [4][stride] the corrector's flux adjoint per unit dt and area
// (qb / V); ZB [5][4][stride] per slot, the adjoint of the face state a cell
// extrapolates to that slot; NG [4][4][stride] per-slot contributions of a
// cell's gradient to the neighbour's state.

// 1. Corrector reverse through the conservative update, per cell:
//   z' = PState(q),  q = Q(z) - dt / V sum_s a_s F_s
// lam (adjoint of z') gives qb; lz = (dQ/dz)^T qb; W = qb / V.  q is rebuilt
// from the stored next state (q = Q(z') up to rounding).  Clears dtb for the
// faces' and the predictor's dt-adjoint sums.
__global__ void __launch_bounds__(256) k_adj_qb(DevMesh M, DevState init, DevState next, const int* __restrict__ wait_in,
                                                Eos eos, DevState lam, DevState lz, double* __restrict__ W,
                                                double* dtb, int* err) {
    const int c = cell_id_a();
    if (c == 0) *dtb = 0.0;
    if (c >= M.n) return;
    if (wait_in[c] > 0) atomicOr(err, kErrRobust);
    const int st = M.stride;
    const Z zn = load_zs(next, c), zc = load_zs(init, c), lb = load_zs(lam, c);
    const Q q = to_q(zn);
    Q qb{0, 0, 0, 0};
    to_z_rev(q, zn, lb, eos, qb);
    Z zcb{0, 0, 0, 0, 0};
    to_q_rev(zc, qb, zcb);
    lz.r[c] = zcb.r; lz.u[c] = zcb.u; lz.v[c] = zcb.v; lz.p[c] = zcb.p; lz.e[c] = zcb.e;
    const double V = M.vol[c];
    W[c] = qb.r / V;
    W[st + c] = qb.mx / V;
    W[2 * st + c] = qb.my / V;
    W[3 * st + c] = qb.E / V;
}

// 2. Corrector faces, once per face: a cell owns its boundary faces and the
// inner / periodic faces of slots 1 and 3 (right, top); the faces of slots 0
// and 2 belong to the neighbour there.  The flux F of a face enters the
// left cell with -dt a / V_c and the right cell with +dt a / V_n, so its
// adjoint is Fb = dt a (W_n - W_c) (boundary: -dt a W_c), and the face adds
// F . Fb / dt to the time step's adjoint.  The HLLC reverse gives the
// adjoints of both face states: the own side goes to ZB[s][c], the
// neighbour's side to ZB[s ^ 1][n] (on a wall / copy face the reflected
// state's adjoint folds back into the own side).
__global__ void __launch_bounds__(256) k_adj_face(DevMesh M, DevState half, DevGrad G, Eos eos, DtRef dref,
                                                  const double* __restrict__ W, double* __restrict__ ZB, double* dtb) {
    const int c = cell_id_a();
    double dtb_loc = 0.0;
    if (c < M.n) {
        const double dt = dt_of(dref);
        const int st = M.stride;
        const int code = M.code[c];
        const Z zch = load_zs(half, c);
        const double ccx = M.cx[c], ccy = M.cy[c];
        double gx[4], gy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gx[k] = G.dx[k * st + c];
            gy[k] = G.dy[k * st + c];
        }
        const double wc0 = W[c], wc1 = W[st + c], wc2 = W[2 * st + c], wc3 = W[3 * st + c];
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            const int t = (code >> (4 * s)) & 7;
            const bool bound = t >= SC_WALL;
            if (!bound && (s & 1) == 0) continue;
            const int so = s * st + c;
            const int nb = M.adj[so];
            const double fx = M.fcx[so], fy = M.fcy[so], a = M.area[so];
            bool bm, bp = false;
            const Z zm = face_state(zch, gx, gy, fx - ccx, fy - ccy, eos, bm);
            Z zp;
            Q Fb;
            if (bound) {
                zp = reflect(zm, s, t == SC_WALL);
                const double k = -dt * a;
                Fb = Q{k * wc0, k * wc1, k * wc2, k * wc3};
            } else {
                double px = M.cx[nb], py = M.cy[nb];
                if (t != SC_INNER) mirror(s, ccx, ccy, fx, fy, px, py);
                double ngx[4], ngy[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    ngx[k] = G.dx[k * st + nb];
                    ngy[k] = G.dy[k * st + nb];
                }
                zp = face_state(load_zs(half, nb), ngx, ngy, fx - px, fy - py, eos, bp);
                const double k = dt * a;
                Fb = Q{k * (W[nb] - wc0), k * (W[st + nb] - wc1), k * (W[2 * st + nb] - wc2),
                       k * (W[3 * st + nb] - wc3)};
            }
            Q lfb = Fb;
            vel_local(s, Fb.mx, Fb.my, lfb.mx, lfb.my);
            Z lmb{0, 0, 0, 0, 0}, lpb{0, 0, 0, 0, 0};
            double fdot = 0.0;
            hllc_rev(to_local(zm, s), to_local(zp, s), lfb, eos, lmb, lpb, &fdot);
            dtb_loc += fdot / dt;
            Z zmb = lmb, zpb = lpb;
            vel_global(s, lmb.u, lmb.v, zmb.u, zmb.v);
            vel_global(s, lpb.u, lpb.v, zpb.u, zpb.v);
            if (bound) {
                zadd(zmb, reflect(zpb, s, t == SC_WALL));
            } else {
                const int o = s ^ 1;
                ZB[(0 * 4 + o) * st + nb] = zpb.r;
                ZB[(1 * 4 + o) * st + nb] = zpb.u;
                ZB[(2 * 4 + o) * st + nb] = zpb.v;
                ZB[(3 * 4 + o) * st + nb] = zpb.p;
                ZB[(4 * 4 + o) * st + nb] = zpb.e;
            }
            ZB[(0 * 4 + s) * st + c] = zmb.r;
            ZB[(1 * 4 + s) * st + c] = zmb.u;
            ZB[(2 * 4 + s) * st + c] = zmb.v;
            ZB[(3 * 4 + s) * st + c] = zmb.p;
            ZB[(4 * 4 + s) * st + c] = zmb.e;
        }
    }
    block_sum_atomic(dtb, dtb_loc);
}

// 3. Per cell: the corrector face states' reverse (half state and gradient
// adjoints from ZB), then the predictor reverse
//   h = PState(Q(z) - dt / (2V) sum_s a_s Flux(zm_s(z, G)))
// adding into lz and writing lG.
__global__ void __launch_bounds__(256) k_adj_pred(DevMesh M, DevState init, DevState half, DevGrad G,
                                                  const int* __restrict__ wait_in, Eos eos, DtRef dref,
                                                  const double* __restrict__ ZB, DevState lz, DevGrad lG,
                                                  double* dtb) {
    const int c = cell_id_a();
    double dtb_loc = 0.0;
    if (c < M.n) {
        const double dt = dt_of(dref);
        const int st = M.stride;
        const double ccx = M.cx[c], ccy = M.cy[c];
        double gx[4], gy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gx[k] = G.dx[k * st + c];
            gy[k] = G.dy[k * st + c];
        }
        Z hb{0, 0, 0, 0, 0};
        double gxb[4] = {0, 0, 0, 0}, gyb[4] = {0, 0, 0, 0};
        {
            const Z zch = load_zs(half, c);
#pragma unroll 1
            for (int s = 0; s < 4; ++s) {
                const int so = s * st + c;
                const double dx = M.fcx[so] - ccx, dy = M.fcy[so] - ccy;
                bool bm;
                const Z zm = face_state(zch, gx, gy, dx, dy, eos, bm);
                const Z zmb{ZB[(0 * 4 + s) * st + c], ZB[(1 * 4 + s) * st + c], ZB[(2 * 4 + s) * st + c],
                            ZB[(3 * 4 + s) * st + c], ZB[(4 * 4 + s) * st + c]};
                face_state_rev(zm, bm, zmb, dx, dy, eos, hb, gxb, gyb);
            }
        }
        const Z zc = load_zs(init, c);
        Z zcb = load_zs(lz, c);
        if (wait_in[c] > 0) {
            zadd(zcb, hb);
        } else {
            double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
#pragma unroll 1
            for (int s = 0; s < 4; ++s) {
                const int so = s * st + c;
                bool bm;
                const Z zm = face_state(zc, gx, gy, M.fcx[so] - ccx, M.fcy[so] - ccy, eos, bm);
                Q f = flux_to_global(phys_flux(to_local(zm, s)), s);
                const double a = M.area[so];
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
            const Z zh = to_z(q, eos);
            if (is_bad(zh, eos)) {
                zadd(zcb, hb);
            } else {
                Q qb{0, 0, 0, 0};
                to_z_rev(q, zh, hb, eos, qb);
                to_q_rev(zc, qb, zcb);
                dtb_loc = -0.5 * (qb.r * F0 + qb.mx * F1 + qb.my * F2 + qb.E * F3) / M.vol[c];
#pragma unroll 1
                for (int s = 0; s < 4; ++s) {
                    const int so = s * st + c;
                    const double dx = M.fcx[so] - ccx, dy = M.fcy[so] - ccy, a = M.area[so];
                    bool bm;
                    const Z zm = face_state(zc, gx, gy, dx, dy, eos, bm);
                    const double w = -coef * a;
                    Q fb{w * qb.r, w * qb.mx, w * qb.my, w * qb.E};
                    Q lfb = fb;
                    vel_local(s, fb.mx, fb.my, lfb.mx, lfb.my);
                    Z lmb{0, 0, 0, 0, 0};
                    phys_flux_rev(to_local(zm, s), lfb, lmb);
                    Z zmb = lmb;
                    vel_global(s, lmb.u, lmb.v, zmb.u, zmb.v);
                    face_state_rev(zm, bm, zmb, dx, dy, eos, zcb, gxb, gyb);
                }
            }
        }
        lz.r[c] = zcb.r; lz.u[c] = zcb.u; lz.v[c] = zcb.v; lz.p[c] = zcb.p; lz.e[c] = zcb.e;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lG.dx[k * st + c] = gxb[k];
            lG.dy[k * st + c] = gyb[k];
        }
    }
    block_sum_atomic(dtb, dtb_loc);
}

// 4. Gradient reverse (LSM + limiting, gradient.h:226-282, 412-474), own part:
// adds the own-state adjoint into lz and writes the neighbour contributions.
template <int LIM>
__global__ void __launch_bounds__(256) k_adj_grad(DevMesh M, DevState init, DevGrad lG, DevState lz,
                                                  double* __restrict__ NG) {
    const int c = cell_id_a();
    if (c >= M.n) return;
    const int st = M.stride;
    const int code = M.code[c];
    const Z zc = load_zs(init, c);
    const double ccx = M.cx[c], ccy = M.cy[c];
    const double zk[4] = {zc.r, zc.u, zc.v, zc.p};
    // forward: slot data, A, unlimited gradient
    double dF[4][4], drx[4], dry[4], dnx[4], dny[4], w[4];
    int typ[4];
    double A00 = 0.0, A01 = 0.0, A10 = 0.0, A11 = 0.0;
    double Fx[4] = {0, 0, 0, 0}, Fy[4] = {0, 0, 0, 0};
    for (int s = 0; s < 4; ++s) {
        const int t = (code >> (4 * s)) & 7;
        typ[s] = t;
        const int so = s * st + c;
        const int nb = M.adj[so];
        Z zn;
        double px, py;
        if (t == SC_INNER || t == SC_PERIODIC) zn = load_zs(init, nb);
        else zn = reflect(zc, s, t == SC_WALL);
        if (t == SC_INNER) {
            px = M.cx[nb];
            py = M.cy[nb];
        } else {
            mirror(s, ccx, ccy, M.fcx[so], M.fcy[so], px, py);
        }
        drx[s] = px - ccx;
        dry[s] = py - ccy;
        dnx[s] = s < 2 ? drx[s] : 0.0;
        dny[s] = s < 2 ? 0.0 : dry[s];
        w[s] = M.w[so];
        const double znk[4] = {zn.r, zn.u, zn.v, zn.p};
#pragma unroll
        for (int k = 0; k < 4; ++k) dF[s][k] = znk[k] - zk[k];
        if (s < 2) {
            const double wd = w[s] * dnx[s];
            A00 = A00 + wd * drx[s];
            A01 = A01 + wd * dry[s];
#pragma unroll
            for (int k = 0; k < 4; ++k) Fx[k] = Fx[k] + (w[s] * dF[s][k]) * dnx[s];
        } else {
            const double wd = w[s] * dny[s];
            A10 = A10 + wd * drx[s];
            A11 = A11 + wd * dry[s];
#pragma unroll
            for (int k = 0; k < 4; ++k) Fy[k] = Fy[k] + (w[s] * dF[s][k]) * dny[s];
        }
    }
    if (fabs(A11) < 1e-8) A11 = 1.0;
    const double det = A11 * A00 + (-A01) * A10;
    const double invdet = 1.0 / det;
    const double B00 = A11 * invdet, B01 = (-A01) * invdet, B10 = (-A10) * invdet, B11 = A00 * invdet;
    double gux[4], guy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gux[k] = B00 * Fx[k] + B01 * Fy[k];
        guy[k] = B10 * Fx[k] + B11 * Fy[k];
    }
    // reverse: G = B L
    double dFb[4][4];
    for (int s = 0; s < 4; ++s)
        for (int k = 0; k < 4; ++k) dFb[s][k] = 0.0;
    double gubx[4] = {0, 0, 0, 0}, guby[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double Gbx = lG.dx[k * st + c], Gby = lG.dy[k * st + c];
        const double Lbx = B00 * Gbx + B10 * Gby, Lby = B01 * Gbx + B11 * Gby;
        for (int s = 0; s < 4; ++s) {
            const double limb = s < 2 ? w[s] * dnx[s] * Lbx : w[s] * dny[s] * Lby;
            const double b = dF[s][k];
            const double du_n = gux[k] * drx[s] + guy[k] * dry[s];
            const double a = 2.0 * du_n - b;
            double pa, pb;
            psi_d<LIM>(a, b, pa, pb);
            const double ab = limb * pa;
            gubx[k] += 2.0 * ab * drx[s];
            guby[k] += 2.0 * ab * dry[s];
            dFb[s][k] += limb * pb - ab;
        }
    }
    // unlimited gradient gu = B F
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double Fbx = B00 * gubx[k] + B10 * guby[k], Fby = B01 * gubx[k] + B11 * guby[k];
        for (int s = 0; s < 4; ++s) dFb[s][k] += s < 2 ? w[s] * dnx[s] * Fbx : w[s] * dny[s] * Fby;
    }
    // dF = zn - zc
    double zcb[4] = {0, 0, 0, 0};
    for (int s = 0; s < 4; ++s) {
        const int t = typ[s];
        double znb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            znb[k] = dFb[s][k];
            zcb[k] -= dFb[s][k];
        }
        if (t == SC_INNER || t == SC_PERIODIC) {
#pragma unroll
            for (int k = 0; k < 4; ++k) NG[(k * 4 + s) * st + c] = znb[k];
        } else {
            const double sg = (t == SC_WALL) ? -1.0 : 1.0;   // reflected normal component
            zcb[0] += znb[0];
            zcb[1] += (s < 2 ? sg : 1.0) * znb[1];
            zcb[2] += (s < 2 ? 1.0 : sg) * znb[2];
            zcb[3] += znb[3];
#pragma unroll
            for (int k = 0; k < 4; ++k) NG[(k * 4 + s) * st + c] = 0.0;
        }
    }
    lz.r[c] += zcb[0];
    lz.u[c] += zcb[1];
    lz.v[c] += zcb[2];
    lz.p[c] += zcb[3];
}

// 5. Gather of the gradient's neighbour contributions and the dt reverse:
// dt = min(CFL tau_min, max_dt), tau_c = h_c / (|v_c| + c_c); the cells
// whose tau equals the minimum receive CFL * dJ/ddt d tau / dz.
__global__ void __launch_bounds__(256) k_adj_final(DevMesh M, DevState init, Eos eos, const double* __restrict__ NG,
                                                   const double* dtb, DtRef dref, DevState lz, DevState lam_out) {
    const int c = cell_id_a();
    if (c >= M.n) return;
    const double tau_min = __longlong_as_double((long long)*dref.slot);
    // dt = smin(CFL tau_min, max_dt): the CFL term is the one taken unless max_dt < CFL tau_min
    const double cfl_active = (*dref.max_dt < dref.cfl * tau_min) ? 0.0 : dref.cfl;
    const int st = M.stride;
    const int code = M.code[c];
    Z zb = load_zs(lz, c);
    for (int s = 0; s < 4; ++s) {
        const int t = (code >> (4 * s)) & 7;
        if (t != SC_INNER && t != SC_PERIODIC) continue;
        const int nb = M.adj[s * st + c], o = s ^ 1;
        zb.r += NG[(0 * 4 + o) * st + nb];
        zb.u += NG[(1 * 4 + o) * st + nb];
        zb.v += NG[(2 * 4 + o) * st + nb];
        zb.p += NG[(3 * 4 + o) * st + nb];
    }
    if (cfl_active != 0.0) {
        const Z z = load_zs(init, c);
        const double vm = sqrt(vsq(z)), cs = eos_c(eos, z.r, z.p);
        const double D = vm + cs;
        const double tau = M.h[c] / D;
        if (tau == tau_min) {
            const double taub = cfl_active * (*dtb);
            const double Db = -taub * tau / D;
            if (vm > 0.0) {
                zb.u += Db * z.u / vm;
                zb.v += Db * z.v / vm;
            }
            eos_c_rev(z.r, z.p, cs, Db, eos, zb.r, zb.p);
        }
    }
    lam_out.r[c] = zb.r;
    lam_out.u[c] = zb.u;
    lam_out.v[c] = zb.v;
    lam_out.p[c] = zb.p;
    lam_out.e[c] = zb.e;
}

// Adjoint seed of J = sum over the window's cells of p^2 V (the integral of
// p^2): lambda_p = 2 p V inside the window, 0 elsewhere.
__global__ void __launch_bounds__(256) k_adj_seed(DevMesh M, DevState z, double x0, double x1, double y0, double y1,
                                                  DevState lam) {
    const int c = cell_id_a();
    if (c >= M.n) return;
    const double x = M.cx[c], y = M.cy[c];
    const bool in = x >= x0 && x <= x1 && y >= y0 && y <= y1;
    lam.r[c] = 0.0;
    lam.u[c] = 0.0;
    lam.v[c] = 0.0;
    lam.e[c] = 0.0;
    lam.p[c] = in ? 2.0 * z.p[c] * M.vol[c] : 0.0;
}

// ---- launchers ----------------------------------------------------------------
static inline dim3 g256(int n) { return dim3((n + 255) / 256); }

hipError_t adj_step(int limiter, const DevMesh& M, const DevState& z, const DevState& znext, const DevState& half,
                    const DevGrad& G, const int* wait, const Eos& eos, const unsigned long long* slot, double cfl,
                    const double* max_dt, const DevState& lam_in, const DevState& lam_out, const DevState& lz,
                    double* W, const DevGrad& lG, double* ZB, double* NG, double* dtb, int* err, hipStream_t s) {
    const DtRef d{slot, cfl, max_dt};
    hipLaunchKernelGGL(k_adj_qb, g256(M.n), dim3(256), 0, s, M, z, znext, wait, eos, lam_in, lz, W, dtb, err);
    hipLaunchKernelGGL(k_adj_face, g256(M.n), dim3(256), 0, s, M, half, G, eos, d, W, ZB, dtb);
    hipLaunchKernelGGL(k_adj_pred, g256(M.n), dim3(256), 0, s, M, z, half, G, wait, eos, d, ZB, lz, lG, dtb);
    switch (limiter) {
        case 0: hipLaunchKernelGGL(k_adj_grad<0>, g256(M.n), dim3(256), 0, s, M, z, lG, lz, NG); break;
        case 1: hipLaunchKernelGGL(k_adj_grad<1>, g256(M.n), dim3(256), 0, s, M, z, lG, lz, NG); break;
        case 2:
        case 12: hipLaunchKernelGGL(k_adj_grad<2>, g256(M.n), dim3(256), 0, s, M, z, lG, lz, NG); break;
        default: hipLaunchKernelGGL(k_adj_grad<3>, g256(M.n), dim3(256), 0, s, M, z, lG, lz, NG); break;
    }
    hipLaunchKernelGGL(k_adj_final, g256(M.n), dim3(256), 0, s, M, z, eos, NG, dtb, d, lz, lam_out);
    return hipGetLastError();
}

hipError_t adj_seed(const DevMesh& M, const DevState& z, const double* win, const DevState& lam, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_seed, g256(M.n), dim3(256), 0, s, M, z, win[0], win[1], win[2], win[3], lam);
    return hipGetLastError();
}

}  // namespace zfv
