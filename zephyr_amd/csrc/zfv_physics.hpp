// Device-side physics of the single-material finite-volume step (gfx950).
//
// 2D state Z = {rho, u, v, p, e}: the reference's smf::PState with v_z = 0
// (math/cfd/models.h:24-28).  Every v_z term of the reference is an exact
// zero (products with 0, sums with +-0), so dropping it changes no value.
// Conserved vector / flux Q = {rho, m_x, m_y, E} (models.h:80-83, 120-123).
//
// Arithmetic order follows the reference's expressions left to right
// (element-wise Eigen array arithmetic).  The library is built with
// -ffp-contract=off, so no multiply-add is fused and +,-,*,/,sqrt are the
// correctly rounded IEEE operations on both the device and the host.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace zfv {

// Slot codes stored 4 bits per slot in the per-cell code word.
enum SlotCode : int {
    SC_UNDEF = 0,     // Boundary::UNDEFINED
    SC_INNER = 1,     // Boundary::INNER
    SC_PERIODIC = 2,  // Boundary::PERIODIC (neighbour state, mirror geometry)
    SC_WALL = 3,      // Boundary::WALL (reflecting boundary value)
    SC_COPY = 4       // Boundary::ZOE / FUNCTION (boundary value = own state)
};

struct Z {
    double r, u, v, p, e;
};
struct Q {
    double r, mx, my, E;
};

struct Eos {
    double g, P0, e0;
};

// std::min / std::max semantics: the first argument wins ties.
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// ---- stiffened-gas EOS (phys/matter/eos/stiffened_gas.cpp:103-175) ----------
__device__ __forceinline__ double eos_p(const Eos& q, double rho, double e) {
    return (q.g - 1.0) * rho * (e - q.e0) - q.g * q.P0;
}
__device__ __forceinline__ double eos_e(const Eos& q, double rho, double p) {
    return q.e0 + (p + q.g * q.P0) / ((q.g - 1.0) * rho);
}
__device__ __forceinline__ double eos_c(const Eos& q, double rho, double p) {
    return sqrt(q.g * (p + q.P0) / rho);
}

// PState::is_bad (math/cfd/models.cpp:58-63).
__device__ __forceinline__ bool is_bad(const Z& z, const Eos& q) {
    if (!isfinite(z.r) || !isfinite(z.p) || !isfinite(z.e)) return true;
    return z.r <= 0.0 || z.p < -q.P0;
}

// |v|^2 with the reference's squaredNorm (v_z^2 = +0 dropped).
__device__ __forceinline__ double vsq(const Z& z) { return z.u * z.u + z.v * z.v; }

// QState(z) (models.cpp:85-89).
__device__ __forceinline__ Q to_q(const Z& z) {
    Q q;
    q.r = z.r;
    q.mx = z.r * z.u;
    q.my = z.r * z.v;
    q.E = z.r * (z.e + 0.5 * vsq(z));
    return q;
}

// PState(q, eos) (models.cpp:27-32).
__device__ __forceinline__ Z to_z(const Q& q, const Eos& eos) {
    Z z;
    z.r = q.r;
    z.u = q.mx / z.r;
    z.v = q.my / z.r;
    z.e = q.E / z.r - 0.5 * vsq(z);
    z.p = eos_p(eos, z.r, z.e);
    return z;
}

// Flux(z), differential flux along the local x axis (models.cpp:131-137).
__device__ __forceinline__ Q phys_flux(const Z& z) {
    Q f;
    double ru = z.r * z.u;
    f.r = ru;
    f.mx = ru * z.u + z.p;
    f.my = ru * z.v;
    f.E = (z.r * (z.e + 0.5 * vsq(z)) + z.p) * z.u;
    return f;
}

// ---- rotations to / from the face frame -----------------------------------
// Rotate::to_local / to_global for the canonical normals of AMR-2D sides
// (rotate.cpp:7-53 with n = -e_x, +e_x, -e_y, +e_y).  The reference's 3x3
// products reduce to these sign/permutation maps with exact zero terms.
__device__ __forceinline__ void vel_local(int side, double u, double v, double& lu, double& lv) {
    switch (side) {
        case 0: lu = -u; lv = -v; break;   // L: n = (-1, 0)
        case 1: lu = u;  lv = v;  break;   // R: n = (+1, 0)
        case 2: lu = -v; lv = u;  break;   // B: n = (0, -1)
        default: lu = v; lv = -u; break;   // T: n = (0, +1)
    }
}
__device__ __forceinline__ void vel_global(int side, double fx, double fy, double& gx, double& gy) {
    switch (side) {
        case 0: gx = -fx; gy = -fy; break;
        case 1: gx = fx;  gy = fy;  break;
        case 2: gx = fy;  gy = -fx; break;
        default: gx = -fy; gy = fx; break;
    }
}
__device__ __forceinline__ Z to_local(const Z& z, int side) {
    Z r = z;
    vel_local(side, z.u, z.v, r.u, r.v);
    return r;
}
__device__ __forceinline__ Q flux_to_global(const Q& f, int side) {
    Q r = f;
    vel_global(side, f.mx, f.my, r.mx, r.my);
    return r;
}

// boundary_value (math/solver/sm_fluid.cpp:69-79): WALL reflects the normal
// velocity, v - 2 (v.n) n, which is exactly a sign flip of one component.
__device__ __forceinline__ Z boundary_value(const Z& z, int side, bool wall) {
    Z r = z;
    if (wall) {
        if (side < 2) r.u = z.u - 2.0 * z.u;
        else          r.v = z.v - 2.0 * z.v;
    }
    return r;
}

// ---- face frames ------------------------------------------------------------
// A slot's frame is its canonical side normal, or, for AMR faces whose
// vertices carry round-off, its stored normal n.  The general formulas are
// the reference's with n_z = 0 and Eigen's order fixed as
// ((a0 b0 + a1 b1) + a2 b2): Rotate (rotate.cpp:7-53) with c, s from
// n.normalized(); boundary_value v - 2 (v.n) n (sm_fluid.cpp:69-79);
// symm_point p + (2 (f - p).n) n (amr_faces.h:197-199); drn = (dr.n) n
// (gradient.h:254).  For a canonical normal they give the same values as the
// side-specialised maps above.
struct Frame {
    int side;
    bool gen;
    double nx, ny;   // stored normal
    double c, s;     // normalised normal
};

__device__ __forceinline__ Frame frame_of(int side, bool gen, double nx, double ny) {
    Frame f;
    f.side = side;
    f.gen = gen;
    f.nx = nx;
    f.ny = ny;
    f.c = nx;
    f.s = ny;
    if (gen) {
        double z = nx * nx + ny * ny;
        if (z > 0.0) {
            double r = sqrt(z);
            f.c = nx / r;
            f.s = ny / r;
        }
    }
    return f;
}

__device__ __forceinline__ Z frame_local(const Frame& f, const Z& z) {
    if (!f.gen) return to_local(z, f.side);
    Z r = z;
    r.u = f.c * z.u + f.s * z.v;
    r.v = (-f.s) * z.u + f.c * z.v;
    return r;
}

__device__ __forceinline__ Q frame_global(const Frame& f, const Q& q) {
    if (!f.gen) return flux_to_global(q, f.side);
    Q r = q;
    r.mx = f.c * q.mx + (-f.s) * q.my;
    r.my = f.s * q.mx + f.c * q.my;
    return r;
}

__device__ __forceinline__ Z frame_boundary(const Frame& f, const Z& z, bool wall) {
    if (!f.gen) return boundary_value(z, f.side, wall);
    Z r = z;
    if (wall) {
        double d = z.u * f.nx + z.v * f.ny;
        r.u = z.u - 2.0 * (f.nx * d);
        r.v = z.v - 2.0 * (f.ny * d);
    }
    return r;
}

__device__ __forceinline__ void frame_symm(const Frame& f, double px, double py, double fx, double fy,
                                           double& sx, double& sy) {
    if (!f.gen) {
        if (f.side < 2) {
            sx = px + 2.0 * (fx - px);
            sy = py;
        } else {
            sx = px;
            sy = py + 2.0 * (fy - py);
        }
        return;
    }
    double d2 = 2.0 * ((fx - px) * f.nx + (fy - py) * f.ny);
    sx = px + d2 * f.nx;
    sy = py + d2 * f.ny;
}

// ---- numerical fluxes in the face frame ------------------------------------
// HLLC::calc_flux, smf (math/cfd/flux/hllc.cpp:23-83).
__device__ __forceinline__ Q flux_hllc(const Z& L, const Z& R, const Eos& eos, bool& nan_out) {
    double cL = eos_c(eos, L.r, L.p);
    double cR = eos_c(eos, R.r, R.p);
    double SL = smin(L.u - cL, R.u - cR);
    double SR = smax(L.u + cL, R.u + cR);
    if (SL >= 0.0) return phys_flux(L);
    if (SR <= 0.0) return phys_flux(R);

    double aL = L.r * (SL - L.u);
    double aR = R.r * (SR - R.u);
    double SC = (L.p - R.p + aR * R.u - aL * L.u) / (aR - aL);
    double rsL = aL / (SL - SC);
    double rsR = aR / (SR - SC);

    double hL = L.e + 0.5 * vsq(L);          // PState::E()
    double hR = R.e + 0.5 * vsq(R);
    double EsL = hL + (SC - L.u) * (SC + L.p / aL);
    double EsR = hR + (SC - R.u) * (SC + R.p / aR);

    // star and side conserved vectors, side fluxes
    double qsL[4] = {rsL, rsL * SC, rsL * L.v, rsL * EsL};
    double qsR[4] = {rsR, rsR * SC, rsR * R.v, rsR * EsR};
    double ruL = L.r * L.u, ruR = R.r * R.u;
    double EL = L.r * hL, ER = R.r * hR;
    double qL[4] = {L.r, ruL, L.r * L.v, EL};
    double qR[4] = {R.r, ruR, R.r * R.v, ER};
    double fL[4] = {ruL, ruL * L.u + L.p, ruL * L.v, (EL + L.p) * L.u};
    double fR[4] = {ruR, ruR * R.u + R.p, ruR * R.v, (ER + R.p) * R.u};

    double aSC = fabs(SC);
    double F[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double t = fL[k] + fR[k];
        t = t + SL * (qsL[k] - qL[k]);
        t = t + aSC * (qsL[k] - qsR[k]);
        t = t + SR * (qsR[k] - qR[k]);
        F[k] = 0.5 * t;
    }
    nan_out |= isnan(F[0]) | isnan(F[1]) | isnan(F[2]) | isnan(F[3]);
    return Q{F[0], F[1], F[2], F[3]};
}

// HLL::calc_flux, smf (math/cfd/flux/hll.cpp:23-57).
__device__ __forceinline__ Q flux_hll(const Z& L, const Z& R, const Eos& eos, bool& nan_out) {
    double cL = eos_c(eos, L.r, L.p);
    double cR = eos_c(eos, R.r, R.p);
    double SL = smin(L.u - cL, smin(R.u - cR, 0.0));
    double SR = smax(L.u + cL, smax(R.u + cR, 0.0));
    double ruL = L.r * L.u, ruR = R.r * R.u;
    double EL = L.r * (L.e + 0.5 * vsq(L)), ER = R.r * (R.e + 0.5 * vsq(R));
    double qL[4] = {L.r, ruL, L.r * L.v, EL};
    double qR[4] = {R.r, ruR, R.r * R.v, ER};
    double fL[4] = {ruL, ruL * L.u + L.p, ruL * L.v, (EL + L.p) * L.u};
    double fR[4] = {ruR, ruR * R.u + R.p, ruR * R.v, (ER + R.p) * R.u};
    double SLSR = SL * SR;
    double den = SR - SL;
    double F[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double t = SR * fL[k] - SL * fR[k] + SLSR * (qR[k] - qL[k]);
        F[k] = t / den;
    }
    nan_out |= isnan(F[0]) | isnan(F[1]) | isnan(F[2]) | isnan(F[3]);
    return Q{F[0], F[1], F[2], F[3]};
}

// Rusanov::calc_flux, smf (math/cfd/flux/rusanov.cpp:13-40); the
// initializer-list min/max keep the first extreme element.
__device__ __forceinline__ Q flux_rusanov(const Z& L, const Z& R, const Eos& eos) {
    double cL = eos_c(eos, L.r, L.p);
    double cR = eos_c(eos, R.r, R.p);
    double SL = L.u - cL;
    double t1 = R.u - cR;
    if (t1 < SL) SL = t1;
    if (0.0 < SL) SL = 0.0;
    double SR = L.u + cL;
    double t2 = R.u + cR;
    if (SR < t2) SR = t2;
    if (SR < 0.0) SR = 0.0;
    double S = smax(fabs(SL), SR);
    double ruL = L.r * L.u, ruR = R.r * R.u;
    double EL = L.r * (L.e + 0.5 * vsq(L)), ER = R.r * (R.e + 0.5 * vsq(R));
    double qL[4] = {L.r, ruL, L.r * L.v, EL};
    double qR[4] = {R.r, ruR, R.r * R.v, ER};
    double fL[4] = {ruL, ruL * L.u + L.p, ruL * L.v, (EL + L.p) * L.u};
    double fR[4] = {ruR, ruR * R.u + R.p, ruR * R.v, (ER + R.p) * R.u};
    double hS = 0.5 * S;
    double F[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k] = 0.5 * (fL[k] + fR[k]) + hS * (qL[k] - qR[k]);
    return Q{F[0], F[1], F[2], F[3]};
}

// ---- exact Riemann solver (math/solver/riemann.cpp:12-460) -------------------
struct RiemannFace {
    double rho, Uf, Pf, U;
    bool fail;
};

__device__ __forceinline__ double rs_fK(double P, double pK, double cK, double gK, double p0K,
                                        double AK, double BK, double GK) {
    if (P >= pK) return (P - pK) * sqrt(AK / (P + BK));
    return 2.0 * cK / (gK - 1.0) * (pow((P + p0K) / (pK + p0K), GK) - 1.0);
}
__device__ __forceinline__ double rs_dfK(double P, double rK, double pK, double cK, double p0K,
                                         double AK, double BK, double GK) {
    if (P >= pK) return (1.0 - (P - pK) / (2.0 * (P + BK))) * sqrt(AK / (P + BK));
    return pow((P + p0K) / (pK + p0K), GK - 1.0) / (rK * cK);
}

__device__ inline RiemannFace riemann_face(double rL, double uL, double pL,
                                           double rR, double uR, double pR,
                                           double g, double p0) {
    RiemannFace out;
    out.fail = false;
    double cL = sqrt(g * (pL + p0) / rL);
    double cR = sqrt(g * (pR + p0) / rR);
    // contact_p (riemann.cpp:118-234)
    double A_L = 2.0 / ((g + 1) * rL), A_R = 2.0 / ((g + 1) * rR);
    double B_L = (g - 1.0) * pL / (g + 1.0) + 2.0 * g * p0 / (g + 1.0);
    double B_R = (g - 1.0) * pR / (g + 1.0) + 2.0 * g * p0 / (g + 1.0);
    double G = (g - 1.0) / (2.0 * g);
    double p_dno = smax(-p0, -p0);
    double f_dno = rs_fK(p_dno, pL, cL, g, p0, A_L, B_L, G) + rs_fK(p_dno, pR, cR, g, p0, A_R, B_R, G) + uR - uL;
    double P, U;
    if (f_dno > 0.0) {
        P = p_dno;
        U = 0.0;
    } else {
        // p_init (riemann.cpp:32-44): every classification branch uses it.
        double rcL = rL * cL, rcR = rR * cR;
        double Pa = (pL * rcR + pR * rcL + (uL - uR) * rcL * rcR) / (rcL + rcR);
        double Pmin = -smin(p0, p0) + 1.0e-2 * smin(pL + p0, pR + p0);
        P = smax(Pmin, Pa);
        double DP = smax(1.0e-5 * fabs(pL - pR), 1.0e-12 * fabs(smax(pL + p0, pR + p0)));
        double fL = 0.0, fR = 0.0, Pn = P;
        int it = 0;
        for (; it < 50; ++it) {
            fL = rs_fK(P, pL, cL, g, p0, A_L, B_L, G);
            fR = rs_fK(P, pR, cR, g, p0, A_R, B_R, G);
            double dfL = rs_dfK(P, rL, pL, cL, p0, A_L, B_L, G);
            double dfR = rs_dfK(P, rR, pR, cR, p0, A_R, B_R, G);
            Pn = P - (fL + fR + uR - uL) / (dfL + dfR);
            if (Pn < p_dno) Pn = (p_dno + P) / 2;
            if (fabs(Pn - P) < DP && it > 3) break;
            P = Pn;
        }
        if (it >= 50 && fabs((P - Pn) / (P + Pn)) > 0.02) out.fail = true;
        U = 0.5 * (uL + uR + fR - fL);
    }
    out.U = U;
    // RiemannSolver::solve sampling at xi = 0 (riemann.cpp:381-459)
    if (U > 0.0) {
        if (pL < P) {
            double D = uL + (P - pL) / (rL * (U - uL));
            if (D > 0.0) { out.rho = rL; out.Uf = uL; out.Pf = pL; }
            else { out.rho = rL * (P - pL) / ((P - pL) - rL * ((U - uL) * (U - uL))); out.Uf = U; out.Pf = P; }
        } else if (uL - cL > 0.0) {
            out.rho = rL; out.Uf = uL; out.Pf = pL;
        } else {
            double cl = cL + 0.5 * (g - 1.0) * (uL - U);
            if (U - cl < 0.0) {
                out.rho = g * (P + p0) / (cl * cl); out.Uf = U; out.Pf = P;
            } else {
                double base = 2.0 / (g + 1.0) + ((g - 1.0) * (uL - 0.0)) / ((g + 1.0) * cL);
                out.rho = rL * pow(base, 2.0 / (g - 1.0));
                out.Uf = ((g - 1.0) * uL + 2.0 * (cL + 0.0)) / (g + 1.0);
                out.Pf = (pL + p0) * pow(base, 2.0 * g / (g - 1.0)) - p0;
            }
        }
    } else {
        if (P > pR) {
            double D = uR + (P - pR) / (rR * (U - uR));
            if (D < 0.0) { out.rho = rR; out.Uf = uR; out.Pf = pR; }
            else { out.rho = rR * (P - pR) / ((P - pR) - rR * ((U - uR) * (U - uR))); out.Uf = U; out.Pf = P; }
        } else if (uR + cR < 0.0) {
            out.rho = rR; out.Uf = uR; out.Pf = pR;
        } else {
            double cr = cR - 0.5 * (g - 1.0) * (uR - U);
            if (U + cr > 0.0) {
                out.rho = g * (P + p0) / (cr * cr); out.Uf = U; out.Pf = P;
            } else {
                double base = 2.0 / (g + 1.0) - ((g - 1.0) * (uR - 0.0)) / ((g + 1.0) * cR);
                out.rho = rR * pow(base, 2.0 / (g - 1.0));
                out.Uf = ((g - 1.0) * uR + 2.0 * (0.0 - cR)) / (g + 1.0);
                out.Pf = (pR + p0) * pow(base, 2.0 * g / (g - 1.0)) - p0;
            }
        }
    }
    return out;
}

// Godunov::calc_flux, smf (math/cfd/flux/godunov.cpp:16-32).
__device__ inline Q flux_godunov(const Z& L, const Z& R, const Eos& eos, bool& fail) {
    RiemannFace s = riemann_face(L.r, L.u, L.p, R.r, R.u, R.p, eos.g, eos.P0);
    fail |= s.fail;
    Z z;
    z.r = s.rho;
    z.u = s.Uf;
    z.v = s.U > 0.0 ? L.v : R.v;
    z.p = s.Pf;
    z.e = eos_e(eos, s.rho, s.Pf);
    return phys_flux(z);
}

// HLLC_M::calc_flux (math/cfd/flux/hllc.cpp:491-572): HLLC with the
// tangential velocity averaged across the contact (V_C) and the matching
// energy correction of the star states.  v_z is zero in 2D: the reference's
// v_y^2 + v_z^2 is v_y^2 and its W_C only feeds the z momentum.
__device__ __forceinline__ Q flux_hllc_m(const Z& L, const Z& R, const Eos& eos, bool& nan_out) {
    double cL = eos_c(eos, L.r, L.p);
    double cR = eos_c(eos, R.r, R.p);
    double SL = smin(L.u - cL, R.u - cR);
    double SR = smax(L.u + cL, R.u + cR);
    if (SL >= 0.0) return phys_flux(L);
    if (SR <= 0.0) return phys_flux(R);
    double aL = L.r * (SL - L.u);
    double aR = R.r * (SR - R.u);
    double SC = (L.p - R.p + aR * R.u - aL * L.u) / (aR - aL);
    double VC = (aR * R.v - aL * L.v) / (aR - aL);
    double rsL = aL / (SL - SC);
    double rsR = aR / (SR - SC);
    double hL = L.e + 0.5 * vsq(L);
    double hR = R.e + 0.5 * vsq(R);
    double EsL = hL + (SC - L.u) * (SC + L.p / aL);
    double EsR = hR + (SC - R.u) * (SC + R.p / aR);
    double vL2 = L.v * L.v, vR2 = R.v * R.v;
    EsL = EsL + 0.5 * ((aR * vR2 - aL * vL2) / (aR - aL) - vL2);
    EsR = EsR + 0.5 * ((aR * vR2 - aL * vL2) / (aR - aL) - vR2);
    double qsL[4] = {rsL, rsL * SC, rsL * VC, rsL * EsL};
    double qsR[4] = {rsR, rsR * SC, rsR * VC, rsR * EsR};
    double ruL = L.r * L.u, ruR = R.r * R.u;
    double EL = L.r * hL, ER = R.r * hR;
    double qL[4] = {L.r, ruL, L.r * L.v, EL};
    double qR[4] = {R.r, ruR, R.r * R.v, ER};
    double fL[4] = {ruL, ruL * L.u + L.p, ruL * L.v, (EL + L.p) * L.u};
    double fR[4] = {ruR, ruR * R.u + R.p, ruR * R.v, (ER + R.p) * R.u};
    double aSC = fabs(SC);
    double F[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double t = fL[k] + fR[k];
        t = t + SL * (qsL[k] - qL[k]);
        t = t + aSC * (qsL[k] - qsR[k]);
        t = t + SR * (qsR[k] - qR[k]);
        F[k] = 0.5 * t;
    }
    nan_out |= isnan(F[0]) | isnan(F[1]) | isnan(F[2]) | isnan(F[3]);
    return Q{F[0], F[1], F[2], F[3]};
}

// HLLC_LM::calc_flux (math/cfd/flux/hllc.cpp:427-488): the outer wave terms
// of HLLC scaled by phi = sin(pi/2 Ma / 0.1) below Ma = 0.1, Ma the larger
// |v_x| / c of the two sides.  sin is the device's (ocml), the reference
// calls glibc's: phi may differ in the last bit.
__device__ __forceinline__ Q flux_hllc_lm(const Z& L, const Z& R, const Eos& eos, bool& nan_out) {
    double cL = eos_c(eos, L.r, L.p);
    double cR = eos_c(eos, R.r, R.p);
    double SL = smin(L.u - cL, R.u - cR);
    double SR = smax(L.u + cL, R.u + cR);
    if (SL >= 0.0) return phys_flux(L);
    if (SR <= 0.0) return phys_flux(R);
    double aL = L.r * (SL - L.u);
    double aR = R.r * (SR - R.u);
    double SC = (L.p - R.p + aR * R.u - aL * L.u) / (aR - aL);
    double rsL = aL / (SL - SC);
    double rsR = aR / (SR - SC);
    double hL = L.e + 0.5 * vsq(L);
    double hR = R.e + 0.5 * vsq(R);
    double EsL = hL + (SC - L.u) * (SC + L.p / aL);
    double EsR = hR + (SC - R.u) * (SC + R.p / aR);
    double qsL[4] = {rsL, rsL * SC, rsL * L.v, rsL * EsL};
    double qsR[4] = {rsR, rsR * SC, rsR * R.v, rsR * EsR};
    double ruL = L.r * L.u, ruR = R.r * R.u;
    double EL = L.r * hL, ER = R.r * hR;
    double qL[4] = {L.r, ruL, L.r * L.v, EL};
    double qR[4] = {R.r, ruR, R.r * R.v, ER};
    double fL[4] = {ruL, ruL * L.u + L.p, ruL * L.v, (EL + L.p) * L.u};
    double fR[4] = {ruR, ruR * R.u + R.p, ruR * R.v, (ER + R.p) * R.u};
    const double Ma = smax(fabs(L.u / cL), fabs(R.u / cR));
    const double xi = Ma / 0.1;
    const double phi = xi > 1.0 ? 1.0 : sin(1.5707963267948966 * xi);
    double aSC = fabs(SC);
    double F[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double t = fL[k] + fR[k];
        t = t + phi * SL * (qsL[k] - qL[k]);
        t = t + aSC * (qsL[k] - qsR[k]);
        t = t + phi * SR * (qsR[k] - qR[k]);
        F[k] = 0.5 * t;
    }
    nan_out |= isnan(F[0]) | isnan(F[1]) | isnan(F[2]) | isnan(F[3]);
    return Q{F[0], F[1], F[2], F[3]};
}

// CIR1 / CIR2 (math/cfd/flux/cir.cpp:16-216): F = (F_L + F_R)/2 +
// |A| (q_L - q_R)/2, |A| = R |Lambda| L at the arithmetic face state (energy
// from (rho, p)).  CIR1 decomposes the 5 x 5 Jacobian (it needs dP/de =
// (gamma - 1) rho); CIR2 carries the pressure as a sixth variable (5 x 6 R,
// 6 x 6 L).  The z velocity (zero in 2D) and the z momentum are carried as
// in the reference's vectors so that every sum has the same terms in the
// same order: |A|(i,j) = sum_k (R(i,k) |lambda_k|) L(k,j), then sum_j.
// cir_flux5 takes the sides as {rho, u, v, w, p, e} and gives all five
// components (the three-dimensional step passes its w; the planar one 0).
template <bool EXT>
__device__ __forceinline__ void cir_flux5(const double* zl, const double* zr, const Eos& eos, double* F) {
    constexpr int N = EXT ? 6 : 5;
    const double rho = 0.5 * (zl[0] + zr[0]), u = 0.5 * (zl[1] + zr[1]), v = 0.5 * (zl[2] + zr[2]);
    const double w = 0.5 * (zl[3] + zr[3]);
    const double p = 0.5 * (zl[4] + zr[4]);
    const double eps = eos_e(eos, rho, p);
    const double u2 = u * u, v2 = v * v, w2 = w * w;
    const double q2 = u2 + v2 + w2;
    const double E = rho * (eps + 0.5 * q2);
    const double c = eos_c(eos, rho, p);
    const double c2 = c * c;
    const double h = (E + p) / rho;
    double lam[N], R[5][N], L[N][N];
    if constexpr (!EXT) {
        const double b = ((eos.g - 1.0) * rho) / rho;
        const double T = b * (q2 - h) + c2;
        const double l5[5] = {fabs(u - c), fabs(u), fabs(u), fabs(u), fabs(u + c)};
        const double r5[5][5] = {{1, 0, 0, 1, 1},
                                 {u - c, 0, 0, u, u + c},
                                 {v, 1, 0, v, v},
                                 {w, 0, 1, w, w},
                                 {h - u * c, v, w, h - c2 / b, h + u * c}};
        const double s2 = 2.0 * c2;
        const double L5[5][5] = {{T + u * c, -b * u - c, -b * v, -b * w, b},
                                 {-2 * v * c2, 0, 2 * c2, 0, 0},
                                 {-2 * w * c2, 0, 0, 2 * c2, 0},
                                 {2 * b * (h - q2), 2 * b * u, 2 * b * v, 2 * b * w, -2 * b},
                                 {T - u * c, -b * u + c, -b * v, -b * w, b}};
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            lam[i] = l5[i];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                R[i][j] = r5[i][j];
                L[i][j] = L5[i][j] / s2;
            }
        }
    } else {
        const double l6[6] = {fabs(u - c), fabs(u), fabs(u), fabs(u), fabs(u + c), fabs(u)};
        const double r6[5][6] = {{1, 0, 0, 0, 1, 1},
                                 {u - c, 0, 0, 0, u + c, u},
                                 {v, 1, 0, 0, v, 0},
                                 {w, 0, 1, 0, w, 0},
                                 {h - u * c, 0, 0, 1, h + u * c, 0}};
        const double L6[6][6] = {{u / (2 * c), -1 / (2 * c), 0, 0, 0, 1 / (2 * c2)},
                                 {0, 0, 1, 0, 0, -v / c2},
                                 {0, 0, 0, 1, 0, -w / c2},
                                 {u2, -u, 0, 0, 1, -h / c2},
                                 {-u / (2 * c), 1 / (2 * c), 0, 0, 0, 1 / (2 * c2)},
                                 {1, 0, 0, 0, 0, -1 / c2}};
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            lam[i] = l6[i];
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                if (i < 5) R[i][j] = r6[i][j];
                L[i][j] = L6[i][j];
            }
        }
    }
    // conserved vectors and physical fluxes
    const double sqL = (zl[1] * zl[1] + zl[2] * zl[2]) + zl[3] * zl[3];
    const double sqR = (zr[1] * zr[1] + zr[2] * zr[2]) + zr[3] * zr[3];
    const double EL = zl[0] * (zl[5] + 0.5 * sqL), ER = zr[0] * (zr[5] + 0.5 * sqR);
    const double qL[5] = {zl[0], zl[0] * zl[1], zl[0] * zl[2], zl[0] * zl[3], EL};
    const double qR[5] = {zr[0], zr[0] * zr[1], zr[0] * zr[2], zr[0] * zr[3], ER};
    const double ruL = zl[0] * zl[1], ruR = zr[0] * zr[1];
    const double fL[5] = {ruL, ruL * zl[1] + zl[4], ruL * zl[2], ruL * zl[3], (EL + zl[4]) * zl[1]};
    const double fR[5] = {ruR, ruR * zr[1] + zr[4], ruR * zr[2], ruR * zr[3], (ER + zr[4]) * zr[1]};
    double dq[N];
#pragma unroll
    for (int k = 0; k < 5; ++k) dq[k] = qL[k] - qR[k];
    if constexpr (EXT) dq[5] = zl[4] - zr[4];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double a = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) a = a + (R[i][k] * lam[k]) * L[k][j];
            t = t + a * dq[j];
        }
        F[i] = 0.5 * (fL[i] + fR[i]) + 0.5 * t;
    }
}

// The planar CIR flux: v_z = 0 on both sides (every z term an exact zero,
// (u u + v v) + 0 0 the reference's squaredNorm), the z momentum dropped.
template <bool EXT>
__device__ __forceinline__ Q flux_cir(const Z& ZL, const Z& ZR, const Eos& eos) {
    const double zl[6] = {ZL.r, ZL.u, ZL.v, 0.0, ZL.p, ZL.e}, zr[6] = {ZR.r, ZR.u, ZR.v, 0.0, ZR.p, ZR.e};
    double F[5];
    cir_flux5<EXT>(zl, zr, eos, F);
    return Q{F[0], F[1], F[2], F[4]};
}

// Fluxes values (math/cfd/flux/num_flux.h:11-21).
template <int FLUX>
__device__ __forceinline__ Q num_flux(const Z& L, const Z& R, const Eos& eos, bool& err) {
    if constexpr (FLUX == 0) return flux_godunov(L, R, eos, err);
    else if constexpr (FLUX == 1) return flux_rusanov(L, R, eos);
    else if constexpr (FLUX == 2) return flux_hll(L, R, eos, err);
    else if constexpr (FLUX == 4) return flux_hllc_m(L, R, eos, err);
    else if constexpr (FLUX == 5) return flux_hllc_lm(L, R, eos, err);
    else if constexpr (FLUX == 6) return flux_cir<false>(L, R, eos);
    else if constexpr (FLUX == 7) return flux_cir<true>(L, R, eos);
    else return flux_hllc(L, R, eos, err);
}

// ---- slope limiters (math/cfd/limiter.cxx:15-153), binary form func2(a, b) ----
__device__ __forceinline__ double lim_r(int kind, double r) {
    switch (kind) {
        case 2: return isnan(r) ? 0.0 : smax(0.0, smin(r, 1.0));
        case 3: return isnan(r) ? 0.0 : smax(0.0, smin(2.0 * r, smin(0.5 * (1.0 + r), 2.0)));
        case 10: return isnan(r) ? 0.0 : smax(0.0, smin(2.0 * r, smin(2.0, (1.0 + 2.0 * r) / 3.0)));
        case 11: return isnan(r) ? 0.0 : smax(0.0, smin(2.0 * r, smin(0.25 + 0.75 * r, 4.0)));
        case 12: return isnan(r) ? 0.0 : smax(0.0, smin(smin(2.0 * r, 1.0), smin(r, 2.0)));
        default: return 0.0;
    }
}

template <int LIM>
__device__ __forceinline__ double limiter2(double a, double b) {
    if constexpr (LIM == 0) return 0.0;
    else if constexpr (LIM == 1) return 1.0;
    else if constexpr (LIM == 2) {   // minmod
        if (a * b <= 0.0) return 0.0;
        return smin(a / b, 1.0);
    } else if constexpr (LIM == 4) {   // van Leer
        if (a * b < 0.0) return 0.0;
        if (a == 0.0) return 0.0;
        return 2.0 * fabs(a) / (fabs(a) + fabs(b));
    } else if constexpr (LIM == 5) {   // van Albada 1
        if (a * b <= 0.0) return 0.0;
        if (b == 0.0) return 1.0;
        return a * (a + b) / (a * a + b * b);
    } else if constexpr (LIM == 6) {   // van Albada 2
        if (a * b <= 0.0) return 0.0;
        const double eps = 10e-6;
        return (2.0 * a * b + eps) / (a * a + b * b + eps);
    } else if constexpr (LIM == 7) {   // CHARM
        if (a * b <= 0.0) return 0.0;
        double s = a + b;
        return a * (3.0 * a + b) / (s * s);
    } else if constexpr (LIM == 8) {   // HCUS
        if (a * b <= 0.0) return 0.0;
        return (3.0 * a) / (a + 2.0 * b);
    } else if constexpr (LIM == 9) {   // HQUICK
        if (a * b <= 0.0) return 0.0;
        return (4.0 * a) / (a + 3.0 * b);
    } else {   // MC, Koren, smart, superbee: func1(a / b)
        return lim_r(LIM, a / b);
    }
}

}  // namespace zfv
