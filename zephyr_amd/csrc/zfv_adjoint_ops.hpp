// Reverse-mode (adjoint) versions of the step's primitives: each *_rev adds
// (d out / d in)^T out_bar into the input adjoints, evaluated at the forward
// inputs; branches take the derivative of the branch the forward takes.
// Used by the adjoint kernels (zfv_adjoint.hip); header-only so the
// primitives can also be checked against finite differences on the host.
#pragma once

#include "zfv_physics.hpp"

namespace zfv {
namespace adj {

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

__device__ __forceinline__ double qdot(const Q& a, const Q& b) { return a.r * b.r + a.mx * b.mx + a.my * b.my + a.E * b.E; }

// HLLC (flux_hllc) reverse: Lb, Rb += (dF/dL)^T Fb, (dF/dR)^T Fb.  With
// fdot != nullptr also stores F . Fb, F the forward flux (the adjoint of a
// scalar multiplying the flux, e.g. the time step).
__device__ __forceinline__ void hllc_rev(const Z& L, const Z& R, const Q& Fb, const Eos& eos, Z& Lb, Z& Rb,
                                         double* fdot = nullptr) {
    const double cL = eos_c(eos, L.r, L.p), cR = eos_c(eos, R.r, R.p);
    const double sl1 = L.u - cL, sl2 = R.u - cR, sr1 = L.u + cL, sr2 = R.u + cR;
    const double SL = smin(sl1, sl2), SR = smax(sr1, sr2);
    double cLb = 0.0, cRb = 0.0;
    if (SL >= 0.0) {
        if (fdot) *fdot = qdot(phys_flux(L), Fb);
        phys_flux_rev(L, Fb, Lb);
        return;
    }
    if (SR <= 0.0) {
        if (fdot) *fdot = qdot(phys_flux(R), Fb);
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
    if (fdot) {
        const double fL[4] = {ruL, ruL * L.u + L.p, ruL * L.v, (EL + L.p) * L.u};
        const double fR[4] = {ruR, ruR * R.u + R.p, ruR * R.v, (ER + R.p) * R.u};
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            d += tb[k] * (fL[k] + fR[k] + SL * (qsL[k] - qL[k]) + aSC * (qsL[k] - qsR[k]) + SR * (qsR[k] - qR[k]));
        *fdot = d;
    }
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


}  // namespace adj
}  // namespace zfv
