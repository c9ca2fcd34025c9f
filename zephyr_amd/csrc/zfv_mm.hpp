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
__device__ void find_rT_rP(const Mix& mx, double rho, double P, const double* b, double T0, const double* rh0,
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
__device__ void get_reT(const Mix& mx, double rho, double P, const double* b, double T0, const double* rh0,
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
__device__ double sound_speed(const Mix& mx, double rho, double P, const double* b, double T0, const double* rh0) {
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

// get_rPT (:199-212): P, T and true densities from (rho, e); a mixture by
// find_rPT_old (:448-533), Newton on (P, T) with the P > P_min guard.
template <int NM>
__device__ void get_rPT(const Mix& mx, double rho, double e_mix, const double* b, double P0, double T0,
                        const double* rh0, double* rh, double& P, double& T) {
    const int idx = frac_index<NM>(b);
    if (idx >= 0) {
        const Mat m = pick<NM>(mx, idx);
        pure_rhos<NM>(rh, idx, rho);
        P = sg_pressure_re(m, rho, e_mix);
        T = sg_temperature_rP(m, rho, P);
        return;
    }
    const double v_mix = 1.0 / rho;
    P = isnan(P0) ? 1.0e5 : P0;
    T = isnan(T0) ? 300.0 : T0;
    double P_min = -INFINITY;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (b[i] > 0.0) P_min = smax(P_min, -mx.m[i].P0);
    if (P < P_min) P = P_min + 1;
    init_rhos<NM>(mx, b, rh0, rh);
    for (int counter = 0; counter < kMaxIter; ++counter) {
        double err = 0.0;
        DPDT v{0.0, 0.0, 0.0}, e{0.0, 0.0, 0.0};
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
__device__ void hllc_mm(const MZ<NM>& zL, const MZ<NM>& zR, double* F, bool& bad) {
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

__device__ __forceinline__ double face_fraction(double a1, double a2) {
    double lo, hi;   // minmax_unit (sections.cpp:42-48)
    if (a1 < a2) {
        lo = smax(0.0, a1);
        hi = smin(a2, 1.0);
    } else {
        lo = smax(0.0, a2);
        hi = smin(a1, 1.0);
    }
    if (3.0 * lo >= hi && lo >= 3.0 * hi - 2.0) return 0.5 * (lo + hi);
    if (hi <= 0.5 || lo * (1.0 - hi) >= (hi - 0.5) * (hi - 0.5)) return -2.0 * lo + 2.0 * sqrt(lo * (lo + hi));
    if (lo >= 0.5 || lo * (1.0 - hi) >= (lo - 0.5) * (lo - 0.5))
        return 3.0 - 2.0 * hi - 2.0 * sqrt((1.0 - hi) * (2.0 - lo - hi));
    if (lo == 0.0 && hi == 1.0) return 0.5;
    return sqrt(lo) / (sqrt(1.0 - hi) + sqrt(lo));
}

}  // namespace mm
}  // namespace zfv
