// Device model of zephyr's multi-material solver (MmFluid,
// math/solver/mm_fluid.{h,cpp}) in its first-order form without composite
// Riemann problems (CrpMode::NONE): stiffened-gas materials with a
// temperature (phys/matter/eos/stiffened_gas.cpp), mass fractions
// (phys/fractions.cpp), the pressure-temperature equilibrium closure of the
// mixture (phys/matter/mixture_pt.cpp, old-style Newton iterations, the
// default) and the material-aware HLLC flux (math/cfd/flux/hllc.cpp:339-421).
//
// Up to four materials; NM (the mixture's size) is a template parameter, so
// the per-material loops unroll.  The velocity has no z component (2D): the
// reference's v_z = 0 terms only add exact zeros.  std::pow(x, 2) is x * x
// (glibc's pow(x, 2.0) rounds identically).
#pragma once

#include "zfv_physics.hpp"

namespace zfv {
namespace mm {

constexpr int kMaxMat = 4;

struct Mat {
    double g, P0, e0, Cv, T0, rho_ref;   // StiffenedGas (stiffened_gas.cpp:11-12), ref density
};

struct Mix {
    int n;
    Mat m[kMaxMat];
};

struct DPDT {
    double val, dP, dT;
};

// ---- one material (stiffened_gas.cpp:103-168) ------------------------------
__device__ __forceinline__ double sg_pressure_re(const Mat& m, double rho, double e) {
    return (m.g - 1.0) * rho * (e - m.e0) - m.g * m.P0;
}
__device__ __forceinline__ double sg_energy_rP(const Mat& m, double rho, double P) {
    return m.e0 + (P + m.g * m.P0) / ((m.g - 1.0) * rho);
}
__device__ __forceinline__ double sg_temperature_rP(const Mat& m, double rho, double P) {
    return m.T0 + (P + m.P0) / ((m.g - 1.0) * m.Cv * rho);
}
__device__ __forceinline__ double sg_sound_speed_rP(const Mat& m, double rho, double P) {
    return sqrt(m.g * (P + m.P0) / rho);
}
__device__ __forceinline__ DPDT sg_volume_PT(const Mat& m, double P, double T) {
    DPDT r;
    r.val = ((m.g - 1.0) * m.Cv * (T - m.T0)) / (P + m.P0);
    r.dP = -r.val / (P + m.P0);
    r.dT = +r.val / (T - m.T0);
    return r;
}
__device__ __forceinline__ DPDT sg_energy_PT(const Mat& m, double P, double T) {
    const double k = (P + m.g * m.P0) / (P + m.P0);
    DPDT r;
    r.val = m.e0 + k * m.Cv * (T - m.T0);
    r.dP = (1.0 - m.g) * m.P0 * m.Cv * (T - m.T0) / ((P + m.P0) * (P + m.P0));
    r.dT = k * m.Cv;
    return r;
}

// ---- mass fractions (fractions.cpp:29-80); has(i) = beta_i > 0 ------------
template <int NM>
__device__ __forceinline__ int frac_index(const double* b) {
    int idx = -1;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (b[i] > 0.0) {
            if (idx >= 0) return -1;
            idx = i;
        }
    return idx;
}
template <int NM>
__device__ __forceinline__ bool frac_pure(const double* b) {
    int k = 0;
#pragma unroll
    for (int i = 0; i < NM; ++i) k += b[i] > 0.0 ? 1 : 0;
    return k == 1;
}
template <int NM>
__device__ __forceinline__ void frac_cutoff(double* b, double eps) {
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        if (b[i] < eps) b[i] = 0.0;
        else if (b[i] > 1.0) b[i] = 1.0;
        sum += b[i];
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) b[i] /= sum;
}
constexpr double kFracMinimal = 0x1.0p-54;   // Fractions::minimal = DBL_EPSILON / 4 (fractions.h:14)

// Material idx of a pure cell, by selects (no dynamic indexing of the
// kernel-argument mixture or of register arrays).
template <int NM>
__device__ __forceinline__ Mat pick(const Mix& mx, int idx) {
    Mat r = mx.m[0];
#pragma unroll
    for (int i = 1; i < NM; ++i)
        if (i == idx) r = mx.m[i];
    return r;
}
template <int NM>
__device__ __forceinline__ void pure_rhos(double* rh, int idx, double rho) {
#pragma unroll
    for (int i = 0; i < NM; ++i) rh[i] = (i == idx) ? rho : NAN;
}

// ---- mixture closure (mixture_pt.cpp:101-533) --------------------------------
constexpr int kMaxIter = 30;
constexpr double kTol = 1.0e-12;

// Initial true densities (init_densities, mixture_pt.cpp:240-256): the
// given ones, or each material's reference density.
template <int NM>
__device__ __forceinline__ void init_rhos(const Mix& mx, const double* b, const double* rh0, double* rh) {
#pragma unroll
    for (int i = 0; i < NM; ++i) rh[i] = b[i] > 0.0 ? (rh0 ? rh0[i] : mx.m[i].rho_ref) : NAN;
}

// find_rT_rP_old (:383-432): T from 1/rho = sum_i beta_i v_i(P, T).
template <int NM>
__device__ __forceinline__ void find_rT_rP(const Mix& mx, double rho, double P, const double* b, double T0, const double* rh0,
                           double* rh, double& T, DPDT& v) {
    const double v_mix = 1.0 / rho;
    T = isnan(T0) ? 300.0 : T0;
    init_rhos<NM>(mx, b, rh0, rh);
    v = DPDT{NAN, NAN, NAN};
    for (int counter = 0; counter < kMaxIter; ++counter) {
        double err = 0.0;
        v = DPDT{0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            if (b[i] > 0.0) {
                const DPDT vi = sg_volume_PT(mx.m[i], P, T);
                v.val += b[i] * vi.val;
                v.dP += b[i] * vi.dP;
                v.dT += b[i] * vi.dT;
                err = smax(err, fabs(1.0 - rh[i] * vi.val));
                rh[i] = 1.0 / vi.val;
            }
        }
        err = smax(err, fabs(1.0 - rho * v.val));
        if (err < kTol) break;
        T += (v_mix - v.val) / v.dT;
    }
}

template <int NM>
__device__ __forceinline__ DPDT mix_energy_PT(const Mix& mx, double P, double T, const double* b) {
    DPDT e{0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (b[i] > 0.0) {
            const DPDT ei = sg_energy_PT(mx.m[i], P, T);
            e.val += b[i] * ei.val;
            e.dP += b[i] * ei.dP;
            e.dT += b[i] * ei.dT;
        }
    return e;
}

// get_reT (:214-227): e, T and true densities from (rho, P); pure material
// in closed form, a mixture through find_reT_rP_old (:434-446).
template <int NM>
__device__ __forceinline__ void get_reT(const Mix& mx, double rho, double P, const double* b, double T0, const double* rh0,
                        double* rh, double& e, double& T) {
    const int idx = frac_index<NM>(b);
    if (idx >= 0) {
        const Mat m = pick<NM>(mx, idx);
        pure_rhos<NM>(rh, idx, rho);
        e = sg_energy_rP(m, rho, P);
        T = sg_temperature_rP(m, rho, P);
        return;
    }
    DPDT v;
    find_rT_rP<NM>(mx, rho, P, b, T0, rh0, rh, T, v);
    e = mix_energy_PT<NM>(mx, P, T, b).val;
}

// sound_speed_rP (:101-114): c^2 = (P / rho^2 - de/drho) / (de/dP) at
// constant (rho, P) derivatives of the equilibrium energy.
template <int NM>
__device__ __forceinline__ double sound_speed(const Mix& mx, double rho, double P, const double* b, double T0, const double* rh0) {
    const int idx = frac_index<NM>(b);
    if (idx >= 0) return sg_sound_speed_rP(pick<NM>(mx, idx), rho, P);
    double rh[NM], T;
    DPDT v;
    find_rT_rP<NM>(mx, rho, P, b, T0, rh0, rh, T, v);
    const double TdR = -(v.val * v.val) / v.dT;
    const double TdP = -v.dP / v.dT;
    const DPDT e1 = mix_energy_PT<NM>(mx, P, T, b);
    const double edR = e1.dT * TdR;
    const double edP = e1.dT * TdP + e1.dP;
    const double c2 = (P / (rho * rho) - edR) / edP;
    return sqrt(c2);
}

template <int NM>
__device__ __forceinline__ void find_rPT(const Mix& mx, double rho, double e_mix, const double* b, double P0, double T0,
                         const double* rh0, double* rh, double& P, double& T, DPDT& v, DPDT& e);

// get_rPT (:199-212): P, T and true densities from (rho, e); a mixture by
// find_rPT_old (:448-533), Newton on (P, T) with the P > P_min guard.
template <int NM>
__device__ __forceinline__ void get_rPT(const Mix& mx, double rho, double e_mix, const double* b, double P0, double T0,
                        const double* rh0, double* rh, double& P, double& T) {
    const int idx = frac_index<NM>(b);
    if (idx >= 0) {
        const Mat m = pick<NM>(mx, idx);
        pure_rhos<NM>(rh, idx, rho);
        P = sg_pressure_re(m, rho, e_mix);
        T = sg_temperature_rP(m, rho, P);
        return;
    }
    DPDT vl, el;
    find_rPT<NM>(mx, rho, e_mix, b, P0, T0, rh0, rh, P, T, vl, el);
}

// find_rPT_old (mixture_pt.cpp:448-533); v, e: the last iteration's sums
// (for the derivatives of sound_speed_re).
template <int NM>
__device__ __forceinline__ void find_rPT(const Mix& mx, double rho, double e_mix, const double* b, double P0, double T0,
                         const double* rh0, double* rh, double& P, double& T, DPDT& v, DPDT& e) {
    const double v_mix = 1.0 / rho;
    P = isnan(P0) ? 1.0e5 : P0;
    T = isnan(T0) ? 300.0 : T0;
    double P_min = -INFINITY;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (b[i] > 0.0) P_min = smax(P_min, -mx.m[i].P0);
    if (P < P_min) P = P_min + 1;
    init_rhos<NM>(mx, b, rh0, rh);
    v = DPDT{NAN, NAN, NAN};
    e = DPDT{NAN, NAN, NAN};
    for (int counter = 0; counter < kMaxIter; ++counter) {
        double err = 0.0;
        v = DPDT{0.0, 0.0, 0.0};
        e = DPDT{0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            if (b[i] > 0.0) {
                const DPDT vi = sg_volume_PT(mx.m[i], P, T);
                v.val += b[i] * vi.val;
                v.dP += b[i] * vi.dP;
                v.dT += b[i] * vi.dT;
                err = smax(err, fabs(1.0 - rh[i] * vi.val));
                rh[i] = 1.0 / vi.val;
                const DPDT ei = sg_energy_PT(mx.m[i], P, T);
                e.val += b[i] * ei.val;
                e.dP += b[i] * ei.dP;
                e.dT += b[i] * ei.dT;
            }
        }
        err = smax(err, fabs(1.0 - rho * v.val));
        const double err_e = fabs(1.0 - e.val / e_mix);
        if (err < kTol && err_e < kTol) break;
        const double inv_D = 1.0 / (v.dT * e.dP - v.dP * e.dT);
        const double dP = inv_D * ((e_mix - e.val) * v.dT - (v_mix - v.val) * e.dT);
        const double dT = inv_D * ((v_mix - v.val) * e.dP - (e_mix - e.val) * v.dP);
        double dP_lim = dP;
        if (P + dP < P_min) dP_lim = (1.0 - 0.1) * P_min + (0.1 - 1.0) * P;
        const double dT_lim = dP != 0.0 ? (dP_lim / dP) * dT : dT;
        P += dP_lim;
        T += dT_lim;
    }
}

// ---- state and HLLC -----------------------------------------------------------
// A cell's primitive state (mmf::PState, models.h:164-283) plus its cached
// sound speed (a function of the state alone, so equal to the value the
// reference recomputes in every flux).
template <int NM>
struct MZ {
    double r, u, v, P, e, T, c;
    double b[NM], rh[NM];
};

template <int NM>
__device__ __forceinline__ double m_E(const MZ<NM>& z) {
    return z.e + 0.5 * (z.u * z.u + z.v * z.v);
}

// Conserved and physical-flux vectors {rho, m_x, m_y, E, rho beta_i}
// (QState(z), Flux(z): models.cpp:413-420, 459-469).
template <int NM>
__device__ __forceinline__ void m_qstate(const MZ<NM>& z, double* q) {
    q[0] = z.r;
    q[1] = z.r * z.u;
    q[2] = z.r * z.v;
    q[3] = z.r * m_E(z);
#pragma unroll
    for (int i = 0; i < NM; ++i) q[4 + i] = z.r * z.b[i];
}
template <int NM>
__device__ __forceinline__ void m_flux(const MZ<NM>& z, double* f) {
    const double ru = z.r * z.u;
    f[0] = ru;
    f[1] = ru * z.u + z.P;
    f[2] = ru * z.v;
    f[3] = (z.r * m_E(z) + z.P) * z.u;
#pragma unroll
    for (int i = 0; i < NM; ++i) f[4 + i] = ru * z.b[i];
}

// HLLC::calc_flux, mmf (hllc.cpp:339-421), in the face frame.
template <int NM>
__device__ __forceinline__ void hllc_mm(const MZ<NM>& zL, const MZ<NM>& zR, double* F, bool& bad) {
    const double c_L = zL.c, c_R = zR.c;
    double S_L = smin(zL.u - c_L, zR.u - c_L);
    double S_R = smax(zL.u + c_R, zR.u + c_R);
    if (S_L >= S_R || frac_index<NM>(zL.b) == frac_index<NM>(zR.b)) {
        S_L = smin(zL.u - c_L, zR.u - c_R);
        S_R = smax(zL.u + c_L, zR.u + c_R);
    }
    if (S_L >= 0.0) {
        m_flux(zL, F);
        return;
    }
    if (S_R <= 0.0) {
        m_flux(zR, F);
        return;
    }
    const double a_L = zL.r * (S_L - zL.u);
    const double a_R = zR.r * (S_R - zR.u);
    double S_C = (zL.P - zR.P + a_R * zR.u - a_L * zL.u) / (a_R - a_L);
    if (S_C < S_L) S_C = S_L + 1.0e-9 * smax(c_L, c_R);
    if (S_C > S_R) S_C = S_R - 1.0e-9 * smax(c_L, c_R);
    const double rho_sL = a_L / (S_L - S_C);
    const double rho_sR = a_R / (S_R - S_C);
    const double E_sL = m_E(zL) + (S_C - zL.u) * (S_C + zL.P / a_L);
    const double E_sR = m_E(zR) + (S_C - zR.u) * (S_C + zR.P / a_R);
    double QsL[4 + NM], QsR[4 + NM], QL[4 + NM], QR[4 + NM], FL[4 + NM], FR[4 + NM];
    QsL[0] = rho_sL;
    QsR[0] = rho_sR;
    QsL[1] = rho_sL * S_C;
    QsL[2] = rho_sL * zL.v;
    QsR[1] = rho_sR * S_C;
    QsR[2] = rho_sR * zR.v;
    QsL[3] = rho_sL * E_sL;
    QsR[3] = rho_sR * E_sR;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        QsL[4 + i] = rho_sL * zL.b[i];
        QsR[4 + i] = rho_sR * zR.b[i];
    }
    m_qstate(zL, QL);
    m_qstate(zR, QR);
    m_flux(zL, FL);
    m_flux(zR, FR);
    const double aSC = fabs(S_C);
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) {
        double t = FL[k] + FR[k];
        t = t + S_L * (QsL[k] - QL[k]);
        t = t + aSC * (QsL[k] - QsR[k]);
        t = t + S_R * (QsR[k] - QR[k]);
        F[k] = 0.5 * t;
        bad |= isnan(F[k]);
    }
}

// Rusanov::calc_flux, mmf (rusanov.cpp:46-83): the largest wave speed
// bound, zero included.
template <int NM>
__device__ void rusanov_mm(const MZ<NM>& zL, const MZ<NM>& zR, double* F, bool& bad) {
    const double c_L = zL.c, c_R = zR.c;
    bad |= isnan(c_L) || isnan(c_R);
    const double S_L = smin(smin(zL.u - c_L, zR.u - c_R), 0.0);
    const double S_R = smax(smax(zL.u + c_L, zR.u + c_R), 0.0);
    const double S = smax(fabs(S_L), S_R);
    double QL[4 + NM], QR[4 + NM], FL[4 + NM], FR[4 + NM];
    m_qstate(zL, QL);
    m_qstate(zR, QR);
    m_flux(zL, FL);
    m_flux(zR, FR);
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) {
        F[k] = 0.5 * (FL[k] + FR[k]) + 0.5 * S * (QL[k] - QR[k]);
        if (k < 4) bad |= !isfinite(F[k]);   // Flux::is_bad (models.cpp:535-541)
    }
}

// HLL::calc_flux, mmf (hll.cpp:125-160): S_L bounded with c_L on both
// sides, S_R with c_R, the Davis bounds if they cross; no upwind branches.
template <int NM>
__device__ void hll_mm(const MZ<NM>& zL, const MZ<NM>& zR, double* F, bool& bad) {
    const double c_L = zL.c, c_R = zR.c;
    double S_L = smin(zL.u - c_L, zR.u - c_L);
    double S_R = smax(zL.u + c_R, zR.u + c_R);
    if (S_L >= S_R) {
        S_L = smin(zL.u - c_L, zR.u - c_R);
        S_R = smax(zL.u + c_L, zR.u + c_R);
    }
    double QL[4 + NM], QR[4 + NM], FL[4 + NM], FR[4 + NM];
    m_qstate(zL, QL);
    m_qstate(zR, QR);
    m_flux(zL, FL);
    m_flux(zR, FR);
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) {
        F[k] = (S_R * FL[k] - S_L * FR[k] + S_L * S_R * (QR[k] - QL[k])) / (S_R - S_L);
        bad |= isnan(F[k]);
    }
}

// PState::is_bad, mmf (models.cpp:393-403).
template <int NM>
__device__ __forceinline__ bool mm_is_bad(const MZ<NM>& z) {
    bool empty = true;   // Fractions::empty (fractions.cpp:91-99)
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (z.b[i] > kFracMinimal) empty = false;
    auto bad = [](double x) { return isinf(x) || isnan(x); };
    return bad(z.r) || bad(z.u) || bad(z.v) || bad(z.P) || bad(z.e) || bad(z.T) || empty || z.r < 0.0;
}

// The flux of the faces without a composite Riemann problem: MmFluid's m_nf
// (HLLC by default; HLL or Rusanov after set_method with CrpMode::NONE),
// numbered as the reference's Fluxes (num_flux.h:11-21).
constexpr int kFluxRusanov = 1, kFluxHll = 2, kFluxHllc = 3;
template <int NM>
__device__ __forceinline__ void plain_flux_mm(int kind, const MZ<NM>& zL, const MZ<NM>& zR, double* F, bool& bad) {
    if (kind == kFluxRusanov) rusanov_mm<NM>(zL, zR, F, bad);
    else if (kind == kFluxHll) hll_mm<NM>(zL, zR, F, bad);
    else hllc_mm<NM>(zL, zR, F, bad);
}

// ---- PLIC formulas (geom/sections.cpp:25-37, 86-104) ---------------------------
__device__ __forceinline__ double quad_find_section(double alpha, double nx, double ny, double A, double B) {
    double a = A * fabs(nx), b = B * fabs(ny);
    if (!(a < b)) {   // sorted (math/funcs.h:32-35)
        const double t = a;
        a = b;
        b = t;
    }
    if (alpha <= 0.5 * a / b) return -0.5 * (a + b) + sqrt(2.0 * alpha * a * b);
    if (alpha < 1.0 - 0.5 * a / b) return (alpha - 0.5) * b;
    return +0.5 * (a + b) - sqrt(2.0 * (1.0 - alpha) * a * b);
}

__device__ __forceinline__ void minmax_unit(double a1, double a2, double& lo, double& hi) {   // sections.cpp:42-48
    if (a1 < a2) {
        lo = smax(0.0, a1);
        hi = smin(a2, 1.0);
    } else {
        lo = smax(0.0, a2);
        hi = smin(a1, 1.0);
    }
}

// The heuristic face fractions of CrpMode V3 / V5 / VS (sections.cpp:52-84).
__device__ __forceinline__ double face_fraction_v3(double a1, double a2) {
    double lo, hi;
    minmax_unit(a1, a2, lo, hi);
    if (lo == 0.0) return 0.0;
    if (hi == 1.0) return 1.0;
    if (a1 + a2 < 1.0) return hi;
    if (a1 + a2 > 1.0) return lo;
    return 0.5 * (a1 + a2);
}
__device__ __forceinline__ double face_fraction_v5(double a1, double a2) {
    double lo, hi;
    minmax_unit(a1, a2, lo, hi);
    if (lo == 0.0) return 0.0;
    if (hi == 1.0) return 1.0;
    return 0.5 * (a1 + a2);
}
__device__ __forceinline__ double f_between(double x, double x1, double x2);
__device__ __forceinline__ double face_fraction_s(double a1, double a2) {
    double lo, hi;
    minmax_unit(a1, a2, lo, hi);
    double a_sig = lo / (1.0 - (hi - lo));
    if (isnan(a_sig)) a_sig = 0.5;
    return f_between(a_sig, lo, hi);
}

__device__ __forceinline__ double face_fraction(double a1, double a2) {
    double lo, hi;
    minmax_unit(a1, a2, lo, hi);
    if (3.0 * lo >= hi && lo >= 3.0 * hi - 2.0) return 0.5 * (lo + hi);
    if (hi <= 0.5 || lo * (1.0 - hi) >= (hi - 0.5) * (hi - 0.5)) return -2.0 * lo + 2.0 * sqrt(lo * (lo + hi));
    if (lo >= 0.5 || lo * (1.0 - hi) >= (lo - 0.5) * (lo - 0.5))
        return 3.0 - 2.0 * hi - 2.0 * sqrt((1.0 - hi) * (2.0 - lo - hi));
    if (lo == 0.0 && hi == 1.0) return 0.5;
    return sqrt(lo) / (sqrt(1.0 - hi) + sqrt(lo));
}

// ---- composite Riemann problems (CrpMode::PLIC, the MmFluid default) --------
// Vectors of K = 4 + NM components {rho, m_x, m_y, E, rho beta_i} in the
// face frame; the reference's m_z = 0 component only carries zeros.

template <int NM>
__device__ __forceinline__ double alpha_of(const MZ<NM>& z, int i) {   // PState::alpha (models.cpp:224-227)
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k)
        if (k == i && !isnan(z.rh[k])) {
            const double x = z.b[k] * z.r / z.rh[k];
            const double lo = (1.0 < x) ? 1.0 : x;
            a = (0.0 < lo) ? lo : 0.0;
        }
    return a;
}

// sound_speed_re of the mixture (mixture_pt.cpp:86-99; pure material
// stiffened_gas.cpp:144-146): c^2 = P.dR + P P.dE / rho^2 with the
// derivatives of find_rPT_old (:525-529), default initial guesses.
template <int NM>
__device__ __forceinline__ double sound_speed_re(const Mix& mx, double rho, double e, const double* b) {
    const int idx = frac_index<NM>(b);
    if (idx >= 0) {
        const Mat m = pick<NM>(mx, idx);
        return sqrt(m.g * (m.g - 1.0) * ((e - m.e0) - m.P0 / rho));
    }
    double rh[NM], P, T;
    DPDT v, ee;
    find_rPT<NM>(mx, rho, e, b, NAN, NAN, nullptr, rh, P, T, v, ee);
    const double inv_D = 1.0 / (v.dT * ee.dP - v.dP * ee.dT);
    const double PdE = inv_D * v.dT;
    const double PdR = inv_D * ee.dT * (v.val * v.val);
    return sqrt(PdR + P * PdE / (rho * rho));
}

template <int NM>
struct Wv {   // mmf::WaveConfig3 (hllc.h:21-27)
    double SL, SC, SR;
    double QsL[4 + NM], FsL[4 + NM], QsR[4 + NM], FsR[4 + NM];
};

// HLLC::wave_config from conserved vectors and fluxes (hllc.cpp:256-335).
template <int NM>
__device__ __forceinline__ void wave_config(const Mix& mx, const double* QL, const double* FL, const double* QR, const double* FR,
                            Wv<NM>& w) {
    constexpr int K = 4 + NM;
    const double uL = QL[1] / QL[0], vL = QL[2] / QL[0];
    const double uR = QR[1] / QR[0], vR = QR[2] / QR[0];
    const double eL = QL[3] / QL[0] - 0.5 * (uL * uL + vL * vL);
    const double eR = QR[3] / QR[0] - 0.5 * (uR * uR + vR * vR);
    double bL[NM], bR[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        bL[i] = QL[4 + i] / QL[0];
        bR[i] = QR[4 + i] / QR[0];
    }
    const double cL = sound_speed_re<NM>(mx, QL[0], eL, bL);
    const double cR = sound_speed_re<NM>(mx, QR[0], eR, bR);
    double SL = smin(uL - cL, uR - cL);
    double SR = smax(uL + cR, uR + cR);
    if (SL >= SR || frac_index<NM>(bL) == frac_index<NM>(bR)) {
        SL = smin(uL - cL, uR - cR);
        SR = smax(uL + cL, uR + cR);
    }
    double aL[K], aR[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        aL[k] = FL[k] - SL * QL[k];
        aR[k] = FR[k] - SR * QR[k];
    }
    double SC = (aR[1] - aL[1]) / (aR[0] - aL[0]);
    if (SC < SL) SC = SL + 1.0e-9 * smax(cL, cR);
    if (SC > SR) SC = SR - 1.0e-9 * smax(cL, cR);
    const double rsL = aL[0] / (SC - SL);
    const double rsR = aR[0] / (SC - SR);
    const double EsL = SC * SC + (aL[3] - SC * aL[1]) / aL[0];
    const double EsR = SC * SC + (aR[3] - SC * aR[1]) / aR[0];
    w.SL = SL;
    w.SC = SC;
    w.SR = SR;
    w.QsL[0] = rsL;
    w.QsR[0] = rsR;
    w.QsL[1] = rsL * SC;
    w.QsL[2] = rsL * vL;
    w.QsR[1] = rsR * SC;
    w.QsR[2] = rsR * vR;
    w.QsL[3] = rsL * EsL;
    w.QsR[3] = rsR * EsR;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        w.QsL[4 + i] = rsL * bL[i];
        w.QsR[4 + i] = rsR * bR[i];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        w.FsL[k] = aL[k] + SL * w.QsL[k];
        w.FsR[k] = aR[k] + SR * w.QsR[k];
    }
}

// Characteristics in the (x, t) plane (crp_flux.cpp:13-55).
struct XT {
    double x, t;
};
struct Chr {
    double x, t, S;
    __device__ double edge_t() const {
        const double ts = (S * t - x) / S;
        return ts < t ? INFINITY : ts;
    }
    __device__ XT cross(const Chr& c) const {
        if (S == c.S) return XT{0.5 * (x + c.x), INFINITY};
        XT p;
        p.t = (c.x - x + S * t - c.S * c.t) / (S - c.S);
        if (p.t <= smax(t, c.t)) return XT{0.5 * (x + c.x), INFINITY};
        p.x = 0.5 * (x + c.x + S * (p.t - t) + c.S * (p.t - c.t));
        return p;
    }
};

template <int NM>
__device__ __forceinline__ void copyk(double* d, const double* s) {
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) d[k] = s[k];
}

// out = c1 ? a : (c2 ? b : c), element by element: a select per component
// keeps the candidate vectors in registers, where a selected pointer would
// make them addressable (stack) memory.
template <int NM>
__device__ __forceinline__ void sel3(bool c1, const double* a, bool c2, const double* b, const double* c,
                                     double* out) {
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) out[k] = c1 ? a[k] : (c2 ? b[k] : c[k]);
}

// (F1 (t1 - t0) - Q1 (x1 - x0) + F2 (...) - ...) / (0 - dt) for two segments.
template <int NM>
__device__ __forceinline__ void total2(const double* F1, const double* Q1, double dt1, double dx1, const double* F2,
                                       const double* Q2, double dt2, double dx2, double dt, double* out) {
#pragma unroll
    for (int k = 0; k < 4 + NM; ++k) out[k] = ((F1[k] * dt1 - Q1[k] * dx1) + (F2[k] * dt2 - Q2[k] * dx2)) / (0.0 - dt);
}

// CrpFlux::classic, ZPP version (crp_flux.cpp:234-489).
template <int NM>
__device__ __forceinline__ void crp_classic(const MZ<NM>& zLA, const MZ<NM>& zLB, const MZ<NM>& zRB, const Mix& mx, double delta,
                            double dt, double* out, bool& err) {
    constexpr int K = 4 + NM;
    double qLB[K], fLB[K], Q_R[K], F_R[K];
    m_qstate(zLB, qLB);
    m_flux(zLB, fLB);
    m_qstate(zRB, Q_R);
    m_flux(zRB, F_R);
    Wv<NM> w0;
    wave_config<NM>(mx, qLB, fLB, Q_R, F_R, w0);
    const Chr C_0{-delta, 0.0, zLA.u};
    const Chr C_0L{0.0, 0.0, w0.SL};
    const XT O1 = C_0.cross(C_0L);
    double qA[K], fA[K];
    m_qstate(zLA, qA);
    m_flux(zLA, fA);
    if (O1.x >= 0.0) {
        const double t2u = C_0.edge_t();
        if (dt <= t2u) {
            copyk<NM>(out, fLB);
            return;
        }
        Wv<NM> w1;
        wave_config<NM>(mx, qA, fA, w0.QsL, w0.FsL, w1);
        const double lb = -O1.x / (dt - O1.t);
        // dt <= O1.t or lb < SL: fA / qA; else by lb against SC and SR
        const bool first = !(dt > O1.t) || lb < w1.SL;
        double F[K], Q[K], F2[K], Q2[K];
        sel3<NM>(lb < w1.SC, w1.FsL, lb < w1.SR, w1.FsR, w0.FsL, F2);
        sel3<NM>(lb < w1.SC, w1.QsL, lb < w1.SR, w1.QsR, w0.QsL, Q2);
        sel3<NM>(first, fA, true, F2, F2, F);
        sel3<NM>(first, qA, true, Q2, Q2, Q);
        total2<NM>(F, Q, O1.t - dt, O1.x - 0.0, fLB, qLB, 0.0 - O1.t, 0.0 - O1.x, dt, out);
        return;
    }
    if (w0.SC >= 0.0) {   // mix to clear
        if (dt <= O1.t && O1.x < 0.0) {
            copyk<NM>(out, w0.FsL);
            return;
        }
        Wv<NM> w1;
        wave_config<NM>(mx, qA, fA, w0.QsL, w0.FsL, w1);
        const Chr C_1R{O1.x, O1.t, w1.SR};
        const double t1 = C_1R.edge_t();
        if (dt <= t1 && O1.x < 0.0) {
            copyk<NM>(out, w0.FsL);
            return;
        }
        const double lb = -O1.x / (dt - O1.t);
        double F[K], Q[K];
        sel3<NM>(lb < w1.SL, fA, lb < w1.SC, w1.FsL, w1.FsR, F);
        sel3<NM>(lb < w1.SL, qA, lb < w1.SC, w1.QsL, w1.QsR, Q);
        total2<NM>(F, Q, O1.t - dt, O1.x - 0.0, w0.FsL, w0.QsL, 0.0 - O1.t, 0.0 - O1.x, dt, out);
        return;
    }
    // clear to mix
    if (O1.t >= dt) {
        copyk<NM>(out, w0.FsR);
        return;
    }
    Wv<NM> w1;
    wave_config<NM>(mx, qA, fA, w0.QsL, w0.FsL, w1);
    const Chr C_0C{0.0, 0.0, w0.SC};
    const Chr C_1R{O1.x, O1.t, w1.SR};
    const XT O2 = C_1R.cross(C_0C);
    if (O2.t > dt || O2.x > 0.0) {
        copyk<NM>(out, w0.FsR);
        return;
    }
    Wv<NM> w2;
    wave_config<NM>(mx, w1.QsR, w1.FsR, w0.QsR, w0.FsR, w2);
    const Chr C_2L{O2.x, O2.t, w2.SL};
    const Chr C_1C{O1.x, O1.t, w1.SC};
    const XT O3 = C_1C.cross(C_2L);
    if (O3.x >= 0.0) {   // crp_flux.cpp:419-421 throws
        err = true;
        copyk<NM>(out, w0.FsR);
        return;
    }
    const Chr C_2R{O2.x, O2.t, w2.SR};
    const double t1 = C_2R.edge_t();
    Wv<NM> w3;
    wave_config<NM>(mx, w1.QsL, w1.FsL, w2.QsL, w2.FsL, w3);
    const Chr C_3R{O3.x, O3.t, w3.SR};
    const double t4 = C_3R.edge_t();
    if (dt <= t4) {
        if (dt <= t1) {
            copyk<NM>(out, w0.FsR);
            return;
        }
        const double lb = -O2.x / (dt - O2.t);
        double F[K], Q[K];
        sel3<NM>(lb < w2.SC, w2.FsL, lb < w2.SR, w2.FsR, w0.FsR, F);
        sel3<NM>(lb < w2.SC, w2.QsL, lb < w2.SR, w2.QsR, w0.QsR, Q);
        total2<NM>(F, Q, O2.t - dt, O2.x - 0.0, w0.FsR, w0.QsR, 0.0 - O2.t, 0.0 - O2.x, dt, out);
        return;
    }
    const double lb = -O3.x / (dt - O3.t);
    double F[K], Q[K];
    sel3<NM>(lb < w3.SL, w1.FsL, lb < w3.SC, w3.FsL, w3.FsR, F);
    sel3<NM>(lb < w3.SL, w1.QsL, lb < w3.SC, w3.QsL, w3.QsR, Q);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double g1 = F[k] * (O3.t - dt) - Q[k] * (O3.x - 0.0);
        const double g2 = w2.FsL[k] * (O2.t - O3.t) - w2.QsL[k] * (O2.x - O3.x);
        const double g3 = w0.FsR[k] * (0.0 - O2.t) - w0.QsR[k] * (0.0 - O2.x);
        out[k] = ((g1 + g2) + g3) / (0.0 - dt);
    }
}

// CrpFlux::inverse (crp_flux.cpp:493-512): the classic problem mirrored in x
// (PState::inverse flips u, Flux::inverse negates all but m_x, models.cpp:219-221, 506-515).
template <int NM>
__device__ __forceinline__ void crp_inverse(const MZ<NM>& zLA, const MZ<NM>& zRA, const MZ<NM>& zRB, const Mix& mx, double delta,
                            double dt, double* out, bool& err) {
    MZ<NM> a = zRB, b = zRA, c = zLA;
    a.u = -a.u;
    b.u = -b.u;
    c.u = -c.u;
    crp_classic<NM>(a, b, c, mx, delta, dt, out, err);
    out[0] = -out[0];
    out[2] = -out[2];
    out[3] = -out[3];
#pragma unroll
    for (int i = 0; i < NM; ++i) out[4 + i] = -out[4 + i];
}

__device__ __forceinline__ double sg_energy_rT(const Mat& m, double rho, double T) {   // stiffened_gas.cpp:131-138
    return m.e0 + m.Cv * (T - m.T0) + m.P0 / rho;
}

// PState::split, ZPP version (models.cpp:261-335): pure material iA and
// the mixture of the others at the same velocity, pressure and temperature;
// both with their sound speeds (sound_speed_rP, as the HLLC of the split
// states computes them).
template <int NM>
__device__ __forceinline__ void m_split(const MZ<NM>& z, const Mix& mx, int iA, MZ<NM>& zA, MZ<NM>& zB) {
    double rhoA = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (i == iA) rhoA = z.rh[i];
    const Mat mA = pick<NM>(mx, iA);
    zA.r = rhoA;
    zA.u = z.u;
    zA.v = z.v;
    zA.P = z.P;
    zA.e = sg_energy_rT(mA, rhoA, z.T);
    zA.T = z.T;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        zA.b[i] = (i == iA) ? 1.0 : 0.0;
        zA.rh[i] = (i == iA) ? rhoA : NAN;
    }
    zA.c = sg_sound_speed_rP(mA, zA.r, zA.P);

    double bA = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (i == iA) bA = z.b[i];
    const double beta_B = 1.0 - bA;
    double alpha[NM];
    zB.u = z.u;
    zB.v = z.v;
    zB.P = z.P;
    zB.T = z.T;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        const bool has = z.b[i] > 0.0;
        zB.rh[i] = (has && i != iA) ? z.rh[i] : NAN;
        zB.b[i] = (has && i != iA) ? z.b[i] / beta_B : 0.0;
        alpha[i] = has ? alpha_of<NM>(z, i) : 0.0;
    }
    frac_cutoff<NM>(zB.b, kFracMinimal);
    double aA = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (i == iA) aA = alpha[i];
    const double alpha_B = 1.0 - aA;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        if (i == iA) alpha[i] = 0.0;
        else if (z.b[i] > 0.0) alpha[i] = alpha[i] / alpha_B;
    }
    frac_cutoff<NM>(alpha, kFracMinimal);
    double rB = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (z.b[i] > 0.0 && i != iA) rB = rB + z.rh[i] * alpha[i];
    zB.r = rB;
    double eB = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (z.b[i] > 0.0 && i != iA) {
            zB.b[i] = alpha[i] * z.rh[i] / rB;
            eB += zB.b[i] * sg_energy_rT(mx.m[i], z.rh[i], z.T);
        }
    zB.e = eB;
    zB.c = sound_speed<NM>(mx, zB.r, zB.P, zB.b, zB.T, zB.rh);
}

// master_component (mm_fluid.cpp:201-226).
template <int NM>
__device__ __forceinline__ int master_component(const MZ<NM>& zm, const MZ<NM>& zp) {
    const double vi = 0.5 * (zm.u + zp.u);
    const MZ<NM>& z = vi > 0.0 ? zm : zp;
    int master = 0;
    double amax = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (z.b[i] > 0.0) {
            const double a = alpha_of<NM>(z, i);
            if (a > amax) {
                amax = a;
                master = i;
            }
        }
    return master;
}

// MmFluid::calc_crp_flux (mm_fluid.cpp:259-337).
template <int NM>
__device__ __forceinline__ void calc_crp_flux(const MZ<NM>& zL, const MZ<NM>& zR, double hL, double hR, int iA, double a_sig,
                              double dt, const Mix& mx, double* F, bool& err) {
    constexpr int K = 4 + NM;
    const double aL = alpha_of<NM>(zL, iA), aR = alpha_of<NM>(zR, iA);
    if ((aL == 0.0 || aL == 1.0) && (aR == 0.0 || aR == 1.0)) {
        hllc_mm<NM>(zL, zR, F, err);
        return;
    }
    const double delta_L = (aL < a_sig ? aL / a_sig : (1.0 - aL) / (1.0 - a_sig)) * hL;
    const double delta_R = (aR < a_sig ? aR / a_sig : (1.0 - aR) / (1.0 - a_sig)) * hR;
    double fA[K], fB[K];
    MZ<NM> xA, xB;
    if (aL == 0.0) {
        m_split<NM>(zR, mx, iA, xA, xB);
        hllc_mm<NM>(zL, xA, fA, err);
        crp_inverse<NM>(zL, xB, xA, mx, delta_R, dt, fB, err);
    } else if (aL == 1.0) {
        m_split<NM>(zR, mx, iA, xA, xB);
        crp_inverse<NM>(zL, xA, xB, mx, delta_R, dt, fA, err);
        hllc_mm<NM>(zL, xB, fB, err);
    } else if (aR == 0.0) {
        m_split<NM>(zL, mx, iA, xA, xB);
        hllc_mm<NM>(xA, zR, fA, err);
        crp_classic<NM>(xA, xB, zR, mx, delta_L, dt, fB, err);
    } else if (aR == 1.0) {
        m_split<NM>(zL, mx, iA, xA, xB);
        crp_classic<NM>(xB, xA, zR, mx, delta_L, dt, fA, err);
        hllc_mm<NM>(xB, zR, fB, err);
    } else {
        MZ<NM> yA, yB;
        m_split<NM>(zL, mx, iA, xA, xB);
        m_split<NM>(zR, mx, iA, yA, yB);
        if (aL < aR) {
            crp_classic<NM>(xB, xA, yA, mx, delta_L, dt, fA, err);
            crp_inverse<NM>(xB, yB, yA, mx, delta_R, dt, fB, err);
        } else {
            crp_inverse<NM>(xA, yA, yB, mx, delta_R, dt, fA, err);
            crp_classic<NM>(xA, xB, yB, mx, delta_L, dt, fB, err);
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) F[k] = a_sig * fA[k] + (1.0 - a_sig) * fB[k];
}

// average_flux (geom/sections.cpp:125-148) and between (math/funcs.h:47-50).
__device__ __forceinline__ double f_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }
__device__ __forceinline__ double f_heav(double x) { return x > 0.0 ? 1.0 : 0.0; }
__device__ __forceinline__ void f_sorted(double a, double b, double& lo, double& hi) {
    if (a < b) {
        lo = a;
        hi = b;
    } else {
        lo = b;
        hi = a;
    }
}
__device__ __forceinline__ double f_between(double x, double x1, double x2) {
    double lo, hi;
    f_sorted(x1, x2, lo, hi);
    return smax(lo, smin(x, hi));
}
__device__ __forceinline__ double average_flux(double alpha, double cs, double CFL) {
    double a1, b1, a2, b2;
    f_sorted(fabs(cs), sqrt(1.0 - cs * cs), a1, b1);
    f_sorted(CFL * fabs(cs), sqrt(1.0 - cs * cs), a2, b2);
    double P1;
    if (fabs(2 * alpha - 1) <= 1.0 - a1 / b1) P1 = (alpha - 0.5) * b1;
    else P1 = f_sign(alpha - 0.5) * (0.5 * (a1 + b1) - sqrt((1.0 - fabs(2.0 * alpha - 1.0)) * a1 * b1));
    const double P2 = P1 + 0.5 * (CFL - 1.0) * cs;
    if (fabs(P2) <= 0.5 * (b2 - a2)) return 0.5 + P2 / b2;
    if (fabs(P2) < 0.5 * (b2 + a2)) {
        const double t = fabs(P2) - 0.5 * (a2 + b2);
        return f_heav(P2) - 0.5 * f_sign(P2) * (t * t) / (a2 * b2);
    }
    return f_heav(P2);
}

}  // namespace mm
}  // namespace zfv
