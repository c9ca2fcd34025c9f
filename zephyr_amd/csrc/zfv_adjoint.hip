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
// took.  Contributions of a cell's faces to its neighbours' inputs are
// written per slot and gathered by the neighbour through the opposite slot
// (on a Cartesian mesh slot s of cell c points to n and slot s ^ 1 of n
// points back to c), so no floating-point atomics are needed except for the
// scalar dt adjoint.
#include "zfv_internal.hpp"

namespace zfv {

namespace {

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

__device__ __forceinline__ void zadd(Z& a, const Z& b) {
    a.r += b.r;
    a.u += b.u;
    a.v += b.v;
    a.p += b.p;
    a.e += b.e;
}

// ---- reverse of the state primitives ----------------------------------------
// to_q: q = {r, r u, r v, r (e + |v|^2 / 2)}
__device__ __forceinline__ void to_q_rev(const Z& z, const Q& qb, Z& zb) {
    const double k = z.e + 0.5 * vsq(z);
    zb.r += qb.r + qb.mx * z.u + qb.my * z.v + qb.E * k;
    zb.u += qb.mx * z.r + qb.E * z.r * z.u;
    zb.v += qb.my * z.r + qb.E * z.r * z.v;
    zb.e += qb.E * z.r;
}

// to_z: r = q.r, u = mx / r, v = my / r, e = E / r - |v|^2 / 2, p = p(r, e)
__device__ __forceinline__ void to_z_rev(const Q& q, const Z& z, const Z& zb, const Eos& eos, Q& qb) {
    double eb = zb.e + zb.p * (eos.g - 1.0) * z.r;
    double rb = zb.r + zb.p * (eos.g - 1.0) * (z.e - eos.e0);
    double ub = zb.u - eb * z.u;
    double vb = zb.v - eb * z.v;
    qb.E += eb / z.r;
    rb -= eb * (q.E / z.r) / z.r;
    qb.mx += ub / z.r;
    rb -= ub * z.u / z.r;
    qb.my += vb / z.r;
    rb -= vb * z.v / z.r;
    qb.r += rb;
}

// e = e0 + (p + g P0) / ((g - 1) r): adds the e adjoint into r and p
__device__ __forceinline__ void eos_e_rev(const Z& z, double eb, const Eos& eos, Z& zb) {
    zb.p += eb / ((eos.g - 1.0) * z.r);
    zb.r -= eb * (z.e - eos.e0) / z.r;
}

// c = sqrt(g (p + P0) / r)
__device__ __forceinline__ void eos_c_rev(double r, double p, double c, double cb, const Eos& eos, double& rb,
                                          double& pb) {
    if (!(c > 0.0)) return;
    const double sb = cb / (2.0 * c);
    pb += sb * eos.g / r;
    rb -= sb * (c * c) / r;
}

// phys_flux (local frame): {r u, r u u + p, r u v, (r (e + |v|^2/2) + p) u}
__device__ __forceinline__ void phys_flux_rev(const Z& z, const Q& fb, Z& zb) {
    const double k = z.e + 0.5 * vsq(z);
    const double Ei = z.r * k;
    zb.r += fb.r * z.u + fb.mx * z.u * z.u + fb.my * z.u * z.v + fb.E * z.u * k;
    zb.u += fb.r * z.r + fb.mx * 2.0 * z.r * z.u + fb.my * z.r * z.v + fb.E * ((Ei + z.p) + z.u * z.r * z.u);
    zb.v += fb.my * z.r * z.u + fb.E * z.u * z.r * z.v;
    zb.p += fb.mx + fb.E * z.u;
    zb.e += fb.E * z.u * z.r;
}

// HLLC (flux_hllc) reverse: Lb, Rb += (dF/dL)^T Fb, (dF/dR)^T Fb.
__device__ void hllc_rev(const Z& L, const Z& R, const Q& Fb, const Eos& eos, Z& Lb, Z& Rb) {
    const double cL = eos_c(eos, L.r, L.p), cR = eos_c(eos, R.r, R.p);
    const double sl1 = L.u - cL, sl2 = R.u - cR, sr1 = L.u + cL, sr2 = R.u + cR;
    const double SL = smin(sl1, sl2), SR = smax(sr1, sr2);
    double cLb = 0.0, cRb = 0.0;
    if (SL >= 0.0) {
        phys_flux_rev(L, Fb, Lb);
        return;
    }
    if (SR <= 0.0) {
        phys_flux_rev(R, Fb, Rb);
        return;
    }
    const double aL = L.r * (SL - L.u), aR = R.r * (SR - R.u);
    const double N = L.p - R.p + aR * R.u - aL * L.u, D = aR - aL;
    const double SC = N / D;
    const double dL = SL - SC, dR = SR - SC;
    const double rsL = aL / dL, rsR = aR / dR;
    const double hL = L.e + 0.5 * vsq(L), hR = R.e + 0.5 * vsq(R);
    const double mL = SC - L.u, kL = SC + L.p / aL, mR = SC - R.u, kR = SC + R.p / aR;
    const double EsL = hL + mL * kL, EsR = hR + mR * kR;
    const double ruL = L.r * L.u, ruR = R.r * R.u;
    const double EL = L.r * hL, ER = R.r * hR;
    const double qsL[4] = {rsL, rsL * SC, rsL * L.v, rsL * EsL};
    const double qsR[4] = {rsR, rsR * SC, rsR * R.v, rsR * EsR};
    const double qL[4] = {L.r, ruL, L.r * L.v, EL};
    const double qR[4] = {R.r, ruR, R.r * R.v, ER};
    const double aSC = fabs(SC);
    const double tb[4] = {0.5 * Fb.r, 0.5 * Fb.mx, 0.5 * Fb.my, 0.5 * Fb.E};
    double SLb = 0.0, SRb = 0.0, SCb = 0.0, aSCb = 0.0;
    double qsLb[4], qsRb[4], qLb[4], qRb[4], fLb[4], fRb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        fLb[k] = tb[k];
        fRb[k] = tb[k];
        SLb += tb[k] * (qsL[k] - qL[k]);
        aSCb += tb[k] * (qsL[k] - qsR[k]);
        SRb += tb[k] * (qsR[k] - qR[k]);
        qsLb[k] = tb[k] * (SL + aSC);
        qsRb[k] = tb[k] * (SR - aSC);
        qLb[k] = -tb[k] * SL;
        qRb[k] = -tb[k] * SR;
    }
    SCb += (SC < 0.0 ? -aSCb : aSCb);
    // side fluxes and conserved vectors
    double ruLb = fLb[0] + fLb[1] * L.u + fLb[2] * L.v + qLb[1];
    double ruRb = fRb[0] + fRb[1] * R.u + fRb[2] * R.v + qRb[1];
    double ELb = fLb[3] * L.u + qLb[3], ERb = fRb[3] * R.u + qRb[3];
    Lb.u += fLb[1] * ruL + fLb[3] * (EL + L.p);
    Rb.u += fRb[1] * ruR + fRb[3] * (ER + R.p);
    Lb.v += fLb[2] * ruL + qLb[2] * L.r;
    Rb.v += fRb[2] * ruR + qRb[2] * R.r;
    Lb.p += fLb[1] + fLb[3] * L.u;
    Rb.p += fRb[1] + fRb[3] * R.u;
    Lb.r += qLb[0] + qLb[2] * L.v;
    Rb.r += qRb[0] + qRb[2] * R.v;
    double hLb = ELb * L.r, hRb = ERb * R.r;
    Lb.r += ELb * hL;
    Rb.r += ERb * hR;
    Lb.r += ruLb * L.u;
    Lb.u += ruLb * L.r;
    Rb.r += ruRb * R.u;
    Rb.u += ruRb * R.r;
    // star states
    double rsLb = qsLb[0] + qsLb[1] * SC + qsLb[2] * L.v + qsLb[3] * EsL;
    double rsRb = qsRb[0] + qsRb[1] * SC + qsRb[2] * R.v + qsRb[3] * EsR;
    SCb += qsLb[1] * rsL + qsRb[1] * rsR;
    Lb.v += qsLb[2] * rsL;
    Rb.v += qsRb[2] * rsR;
    const double EsLb = qsLb[3] * rsL, EsRb = qsRb[3] * rsR;
    hLb += EsLb;
    hRb += EsRb;
    double aLb = 0.0, aRb = 0.0;
    {
        const double mb = EsLb * kL, kb = EsLb * mL;
        SCb += mb + kb;
        Lb.u -= mb;
        Lb.p += kb / aL;
        aLb -= kb * L.p / (aL * aL);
    }
    {
        const double mb = EsRb * kR, kb = EsRb * mR;
        SCb += mb + kb;
        Rb.u -= mb;
        Rb.p += kb / aR;
        aRb -= kb * R.p / (aR * aR);
    }
    Lb.e += hLb;
    Lb.u += hLb * L.u;
    Lb.v += hLb * L.v;
    Rb.e += hRb;
    Rb.u += hRb * R.u;
    Rb.v += hRb * R.v;
    // rs = a / (S - SC)
    aLb += rsLb / dL;
    {
        const double db = -rsLb * rsL / dL;
        SLb += db;
        SCb -= db;
    }
    aRb += rsRb / dR;
    {
        const double db = -rsRb * rsR / dR;
        SRb += db;
        SCb -= db;
    }
    // SC = N / D
    {
        const double Nb = SCb / D, Db = -SCb * SC / D;
        Lb.p += Nb;
        Rb.p -= Nb;
        aRb += Nb * R.u + Db;
        Rb.u += Nb * aR;
        aLb += -Nb * L.u - Db;
        Lb.u -= Nb * aL;
    }
    // a = r (S - u)
    Lb.r += aLb * (SL - L.u);
    SLb += aLb * L.r;
    Lb.u -= aLb * L.r;
    Rb.r += aRb * (SR - R.u);
    SRb += aRb * R.r;
    Rb.u -= aRb * R.r;
    // SL = min(uL - cL, uR - cR), SR = max(uL + cL, uR + cR) (smin / smax choice)
    if (sl2 < sl1) {
        Rb.u += SLb;
        cRb -= SLb;
    } else {
        Lb.u += SLb;
        cLb -= SLb;
    }
    if (sr1 < sr2) {
        Rb.u += SRb;
        cRb += SRb;
    } else {
        Lb.u += SRb;
        cLb += SRb;
    }
    eos_c_rev(L.r, L.p, cL, cLb, eos, Lb.r, Lb.p);
    eos_c_rev(R.r, R.p, cR, cRb, eos, Rb.r, Rb.p);
}

// ---- forward pieces recomputed in the reverse sweep --------------------------
__device__ __forceinline__ Z load_zs(const DevState& s, int c) { return Z{s.r[c], s.u[c], s.v[c], s.p[c], s.e[c]}; }

__device__ __forceinline__ Z extrap_a(const Z& z, const double* gx, const double* gy, double dx, double dy) {
    Z r;
    r.r = z.r + (gx[0] * dx + gy[0] * dy);
    r.u = z.u + (gx[1] * dx + gy[1] * dy);
    r.v = z.v + (gx[2] * dx + gy[2] * dy);
    r.p = z.p + (gx[3] * dx + gy[3] * dy);
    r.e = 0.0;
    return r;
}

// Mirror point of a boundary / periodic slot (symm_point, canonical normal).
__device__ __forceinline__ void mirror(int side, double ccx, double ccy, double fx, double fy, double& px,
                                       double& py) {
    const Frame f = frame_of(side, false, 0.0, 0.0);
    frame_symm(f, ccx, ccy, fx, fy, px, py);
}

// Reverse of a face extrapolation z_face = extrap(z, G, dx, dy) followed by
// e = eos_e(r, p) and the bad-state fallback to z.  zb: adjoint of the face
// state (5); adds into the base state's adjoint (5) and gradient adjoints.
__device__ __forceinline__ void face_state_rev(const Z& zf, bool bad, const Z& zfb, double dx, double dy,
                                               const Eos& eos, Z& zbase_b, double* gxb, double* gyb) {
    if (bad) {
        zadd(zbase_b, zfb);
        return;
    }
    Z t = zfb;
    t.e = 0.0;
    eos_e_rev(zf, zfb.e, eos, t);
    const double tk[4] = {t.r, t.u, t.v, t.p};
    zbase_b.r += t.r;
    zbase_b.u += t.u;
    zbase_b.v += t.v;
    zbase_b.p += t.p;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gxb[k] += tk[k] * dx;
        gyb[k] += tk[k] * dy;
    }
}

__device__ __forceinline__ Z face_state(const Z& base, const double* gx, const double* gy, double dx, double dy,
                                        const Eos& eos, bool& bad) {
    Z zf = extrap_a(base, gx, gy, dx, dy);
    zf.e = eos_e(eos, zf.r, zf.p);
    bad = is_bad(zf, eos);
    return bad ? base : zf;
}

// wall reflection of the normal velocity component (its own transpose)
__device__ __forceinline__ Z reflect(const Z& z, int side, bool wall) {
    Z r = z;
    if (wall) {
        if (side < 2) r.u = -z.u;
        else r.v = -z.v;
    }
    return r;
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
// Buffers: NB [13][4][stride] per-slot contributions of a cell's corrector
// faces to the neighbour's half state (5) and gradient (8); NG [4][4][stride]
// per-slot contributions of a cell's gradient to the neighbour's state.

// 1. Corrector reverse, own part.  lam: adjoint of z^{n+1}; writes the
// adjoint of z through Q(z), of the own half state and gradient, the
// neighbour contributions, and adds dJ/ddt.
__global__ void __launch_bounds__(256) k_adj_corr(DevMesh M, DevState init, DevState half, DevGrad G,
                                                  const int* __restrict__ wait_in, Eos eos, DtRef dref,
                                                  DevState lam, DevState lz, DevState lh, DevGrad lG,
                                                  double* __restrict__ NB, double* dtb, int* err) {
    const int c = cell_id_a();
    const double dt = dt_of(dref);
    double dtb_loc = 0.0;
    if (c < M.n) {
        const int st = M.stride;
        const Z zc = load_zs(init, c), zch = load_zs(half, c);
        const Z lb = load_zs(lam, c);
        Z zcb{0, 0, 0, 0, 0}, zchb{0, 0, 0, 0, 0};
        double gxb[4] = {0, 0, 0, 0}, gyb[4] = {0, 0, 0, 0};
        if (wait_in[c] > 0) atomicOr(err, 4);
        const double ccx = M.cx[c], ccy = M.cy[c];
        double gx[4], gy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gx[k] = G.dx[k * st + c];
            gy[k] = G.dy[k * st + c];
        }
        // forward face fluxes (needed for q and dt adjoint)
        const int code = M.code[c];
        Q Fs[4];
        double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
        for (int s = 0; s < 4; ++s) {
            const int t = (code >> (4 * s)) & 7;
            const int so = s * st + c;
            const int nb = M.adj[so];
            const double fx = M.fcx[so], fy = M.fcy[so], a = M.area[so];
            bool bm, bp = false;
            const Z zm = face_state(zch, gx, gy, fx - ccx, fy - ccy, eos, bm);
            Z zp;
            if (t >= SC_WALL) {
                zp = reflect(zm, s, t == SC_WALL);
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
            }
            bool nanf = false;
            Q f = flux_hllc(to_local(zm, s), to_local(zp, s), eos, nanf);
            f = flux_to_global(f, s);
            Fs[s] = Q{f.r * a, f.mx * a, f.my * a, f.E * a};
            F0 = F0 + Fs[s].r;
            F1 = F1 + Fs[s].mx;
            F2 = F2 + Fs[s].my;
            F3 = F3 + Fs[s].E;
        }
        Q q = to_q(zc);
        const double coef = dt / M.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        const Z zn = to_z(q, eos);
        if (is_bad(zn, eos)) {   // the cell is isolated from here on: not differentiated
            atomicOr(err, 4);
            zadd(zcb, lb);
            for (int s = 0; s < 4; ++s)
                for (int k = 0; k < 13; ++k) NB[(k * 4 + s) * st + c] = 0.0;
        } else {
            Q qb{0, 0, 0, 0};
            to_z_rev(q, zn, lb, eos, qb);
            to_q_rev(zc, qb, zcb);
            dtb_loc = -(qb.r * F0 + qb.mx * F1 + qb.my * F2 + qb.E * F3) / M.vol[c];
            const Q Fb{-coef * qb.r, -coef * qb.mx, -coef * qb.my, -coef * qb.E};
            for (int s = 0; s < 4; ++s) {
                const int t = (code >> (4 * s)) & 7;
                const int so = s * st + c;
                const int nb = M.adj[so];
                const double fx = M.fcx[so], fy = M.fcy[so], a = M.area[so];
                bool bm, bp = false;
                const Z zm = face_state(zch, gx, gy, fx - ccx, fy - ccy, eos, bm);
                Z zp, hn{0, 0, 0, 0, 0};
                double px = 0.0, py = 0.0;
                if (t >= SC_WALL) {
                    zp = reflect(zm, s, t == SC_WALL);
                } else {
                    px = M.cx[nb];
                    py = M.cy[nb];
                    if (t != SC_INNER) mirror(s, ccx, ccy, fx, fy, px, py);
                    hn = load_zs(half, nb);
                    double ngx[4], ngy[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        ngx[k] = G.dx[k * st + nb];
                        ngy[k] = G.dy[k * st + nb];
                    }
                    zp = face_state(hn, ngx, ngy, fx - px, fy - py, eos, bp);
                }
                // flux adjoint: global -> local (transpose of flux_to_global)
                Q fb{Fb.r * a, Fb.mx * a, Fb.my * a, Fb.E * a};
                Q lfb = fb;
                vel_local(s, fb.mx, fb.my, lfb.mx, lfb.my);
                Z lmb{0, 0, 0, 0, 0}, lpb{0, 0, 0, 0, 0};
                hllc_rev(to_local(zm, s), to_local(zp, s), lfb, eos, lmb, lpb);
                Z zmb = lmb, zpb = lpb;
                vel_global(s, lmb.u, lmb.v, zmb.u, zmb.v);
                vel_global(s, lpb.u, lpb.v, zpb.u, zpb.v);
                double nbh[13];
#pragma unroll
                for (int k = 0; k < 13; ++k) nbh[k] = 0.0;
                if (t >= SC_WALL) {
                    zadd(zmb, reflect(zpb, s, t == SC_WALL));
                } else {
                    Z hnb{0, 0, 0, 0, 0};
                    double ngxb[4] = {0, 0, 0, 0}, ngyb[4] = {0, 0, 0, 0};
                    face_state_rev(zp, bp, zpb, fx - px, fy - py, eos, hnb, ngxb, ngyb);
                    nbh[0] = hnb.r; nbh[1] = hnb.u; nbh[2] = hnb.v; nbh[3] = hnb.p; nbh[4] = hnb.e;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        nbh[5 + k] = ngxb[k];
                        nbh[9 + k] = ngyb[k];
                    }
                }
#pragma unroll
                for (int k = 0; k < 13; ++k) NB[(k * 4 + s) * st + c] = nbh[k];
                face_state_rev(zm, bm, zmb, fx - ccx, fy - ccy, eos, zchb, gxb, gyb);
            }
        }
        lz.r[c] = zcb.r; lz.u[c] = zcb.u; lz.v[c] = zcb.v; lz.p[c] = zcb.p; lz.e[c] = zcb.e;
        lh.r[c] = zchb.r; lh.u[c] = zchb.u; lh.v[c] = zchb.v; lh.p[c] = zchb.p; lh.e[c] = zchb.e;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lG.dx[k * st + c] = gxb[k];
            lG.dy[k * st + c] = gyb[k];
        }
    }
    block_sum_atomic(dtb, dtb_loc);
}

// 2. Gather of the corrector's neighbour contributions, then the predictor
// reverse (own cell).  Adds into lz and lG.
__global__ void __launch_bounds__(256) k_adj_pred(DevMesh M, DevState init, DevState half, DevGrad G,
                                                  const int* __restrict__ wait_in, Eos eos, DtRef dref,
                                                  DevState lz, DevState lh, DevGrad lG, const double* __restrict__ NB,
                                                  double* dtb) {
    const int c = cell_id_a();
    const double dt = dt_of(dref);
    double dtb_loc = 0.0;
    if (c < M.n) {
        const int st = M.stride;
        const int code = M.code[c];
        Z hb = load_zs(lh, c);
        double gxb[4], gyb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gxb[k] = lG.dx[k * st + c];
            gyb[k] = lG.dy[k * st + c];
        }
        for (int s = 0; s < 4; ++s) {
            const int t = (code >> (4 * s)) & 7;
            if (t != SC_INNER && t != SC_PERIODIC) continue;
            const int nb = M.adj[s * st + c], o = s ^ 1;
            hb.r += NB[(0 * 4 + o) * st + nb];
            hb.u += NB[(1 * 4 + o) * st + nb];
            hb.v += NB[(2 * 4 + o) * st + nb];
            hb.p += NB[(3 * 4 + o) * st + nb];
            hb.e += NB[(4 * 4 + o) * st + nb];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                gxb[k] += NB[((5 + k) * 4 + o) * st + nb];
                gyb[k] += NB[((9 + k) * 4 + o) * st + nb];
            }
        }
        // predictor reverse: h = PState(Q(z) - dt/(2V) sum Flux(zm) S)
        const Z zc = load_zs(init, c);
        Z zcb = load_zs(lz, c);
        if (wait_in[c] > 0) {
            zadd(zcb, hb);
        } else {
            const double ccx = M.cx[c], ccy = M.cy[c];
            double gx[4], gy[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                gx[k] = G.dx[k * st + c];
                gy[k] = G.dy[k * st + c];
            }
            double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
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

// Limiter psi(a, b) = phi(a, b) b and its partial derivatives (MC, minmod).
template <int LIM>
__device__ __forceinline__ void psi_d(double a, double b, double& pa, double& pb) {
    if (LIM == 0) {
        pa = 0.0;
        pb = 0.0;
        return;
    }
    if (LIM == 1) {
        pa = 0.0;
        pb = 1.0;
        return;
    }
    double phi, dphi, r;
    if (LIM == 2) {   // minmod: ab <= 0 -> 0, else smin(a / b, 1)
        if (a * b <= 0.0) {
            pa = 0.0;
            pb = 0.0;
            return;
        }
        r = a / b;
        const bool one = 1.0 < r;
        phi = one ? 1.0 : r;
        dphi = one ? 0.0 : 1.0;
    } else {   // MC: isnan(r) ? 0 : smax(0, smin(2 r, smin((1 + r) / 2, 2)))
        r = a / b;
        if (isnan(r)) {
            pa = 0.0;
            pb = 0.0;
            return;
        }
        const double half = 0.5 * (1.0 + r);
        const bool sat = 2.0 < half;
        const double inner = sat ? 2.0 : half;
        const bool lin = !(inner < 2.0 * r);
        const double mid = lin ? 2.0 * r : inner;
        if (!(0.0 < mid)) {
            pa = 0.0;
            pb = 0.0;
            return;
        }
        phi = mid;
        dphi = lin ? 2.0 : (sat ? 0.0 : 0.5);
    }
    pa = dphi;
    pb = (dphi == 0.0) ? phi : phi - dphi * r;
}

// 3. Gradient reverse (LSM + limiting, gradient.h:226-282, 412-474), own part:
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

// 4. Gather of the gradient's neighbour contributions and the dt reverse:
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

hipError_t adj_step(int limiter, const DevMesh& M, const DevState& z, const DevState& half, const DevGrad& G,
                    const int* wait, const Eos& eos, const unsigned long long* slot, double cfl,
                    const double* max_dt, const DevState& lam_in, const DevState& lam_out, const DevState& lz,
                    const DevState& lh, const DevGrad& lG, double* NB, double* NG, double* dtb, int* err,
                    hipStream_t s) {
    const DtRef d{slot, cfl, max_dt};
    hipLaunchKernelGGL(k_adj_corr, g256(M.n), dim3(256), 0, s, M, z, half, G, wait, eos, d, lam_in, lz, lh, lG, NB,
                       dtb, err);
    hipLaunchKernelGGL(k_adj_pred, g256(M.n), dim3(256), 0, s, M, z, half, G, wait, eos, d, lz, lh, lG, NB, dtb);
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
