// ORACLE — test infrastructure only (included at the end of oracle.cpp).
// CPU restatement of zephyr's multi-material solver MmFluid in its
// first-order form without composite Riemann problems (CrpMode::NONE: the
// material-aware HLLC flux on every face), with the PT-equilibrium mixture
// closure of stiffened gases and the PLIC interface reconstruction.  Each
// function cites the reference file:line it restates.  Operation order
// follows the reference's expressions under -ffp-contract=off; std::pow(x, 2)
// is written x * x (glibc's pow(x, 2.0) equals x * x bit for bit: checked on
// 5e7 random arguments over 200 binades).

#include <limits>

namespace orcmm {

constexpr int kMaxMat = 4;   // Fractions::max_size (phys/fractions.h:14)

// StiffenedGas (phys/matter/eos/stiffened_gas.cpp:11-175), one material.
struct SgMat {
    double g, P0, e0, Cv, T0, rho_ref;
};

struct DPDT { double val, dP, dT; };   // dPdT (phys/matter/eos/eos.h:37-47)

inline double sg_pressure_re(const SgMat& m, double rho, double e) {   // :104-111
    return (m.g - 1.0) * rho * (e - m.e0) - m.g * m.P0;
}
inline double sg_energy_rP(const SgMat& m, double rho, double P) {   // :122-129
    return m.e0 + (P + m.g * m.P0) / ((m.g - 1.0) * rho);
}
inline double sg_energy_rT(const SgMat& m, double rho, double T) {   // :131-138
    return m.e0 + m.Cv * (T - m.T0) + m.P0 / rho;
}
inline double sg_temperature_rP(const SgMat& m, double rho, double P) {   // :140-142
    return m.T0 + (P + m.P0) / ((m.g - 1.0) * m.Cv * rho);
}
inline double sg_sound_speed_rP(const SgMat& m, double rho, double P) {   // :148-150
    return std::sqrt(m.g * (P + m.P0) / rho);
}
inline DPDT sg_volume_PT(const SgMat& m, double P, double T) {   // :152-159
    DPDT r;
    r.val = ((m.g - 1.0) * m.Cv * (T - m.T0)) / (P + m.P0);
    r.dP = -r.val / (P + m.P0);
    r.dT = +r.val / (T - m.T0);
    return r;
}
inline DPDT sg_energy_PT(const SgMat& m, double P, double T) {   // :161-168
    DPDT r;
    r.val = m.e0 + ((P + m.g * m.P0) / (P + m.P0)) * m.Cv * (T - m.T0);
    r.dP = (1.0 - m.g) * m.P0 * m.Cv * (T - m.T0) / ((P + m.P0) * (P + m.P0));
    r.dT = ((P + m.g * m.P0) / (P + m.P0)) * m.Cv;
    return r;
}

// Fractions (phys/fractions.h:12-105, fractions.cpp:29-110).
struct Frac {
    double a[kMaxMat] = {0.0, 0.0, 0.0, 0.0};
    bool has(int i) const { return a[i] > 0.0; }
    int index() const {   // fractions.cpp:49-61
        int idx = -1;
        for (int i = 0; i < kMaxMat; ++i)
            if (has(i)) {
                if (idx < 0) idx = i;
                else return -1;
            }
        return idx;
    }
    bool is_pure() const {   // fractions.cpp:29-41
        bool pure = false;
        for (int i = 0; i < kMaxMat; ++i)
            if (has(i)) {
                if (!pure) pure = true;
                else return false;
            }
        return pure;
    }
    void cutoff(double eps) {   // fractions.cpp:67-80
        double sum = 0.0;
        for (int i = 0; i < kMaxMat; ++i) {
            if (a[i] < eps) a[i] = 0.0;
            else if (a[i] > 1.0) a[i] = 1.0;
            sum += a[i];
        }
        for (int i = 0; i < kMaxMat; ++i) a[i] /= sum;
    }
    void normalize() { cutoff(std::numeric_limits<double>::epsilon() / 4); }   // :63-65
};

// MixturePT (phys/matter/mixture_pt.{h,cpp}), old-style iterations (the
// default, mixture_pt.h:33).
struct Mixture {
    int n = 0;
    SgMat m[kMaxMat];
};

struct MixOpt {   // MixOptions (mixture_pt.h:12-27)
    bool deriv = false;
    double P0 = NAN, T0 = NAN;
    const double* rhos = nullptr;   // empty span = nullptr
};

constexpr int kMaxIter = 30;      // mixture_pt.cpp:298
constexpr double kEps = 1.0e-12;  // epsilon_v, epsilon_e

// init_densities (mixture_pt.cpp:240-256).
inline void init_densities(const Mixture& mx, const Frac& b, const MixOpt& o, double* rhos) {
    for (int i = 0; i < kMaxMat; ++i) rhos[i] = NAN;
    for (int i = 0; i < mx.n; ++i)
        if (b.has(i)) rhos[i] = o.rhos ? o.rhos[i] : mx.m[i].rho_ref;
}

inline double min_pressure(const Mixture& mx, const Frac& b) {   // mixture_pt.cpp:171-179
    double P_min = -std::numeric_limits<double>::infinity();
    for (int i = 0; i < mx.n; ++i)
        if (b.has(i)) P_min = smax(P_min, -mx.m[i].P0);
    return P_min;
}

// find_rT_rP_old (mixture_pt.cpp:383-432): T from v_mix = sum beta_i v_i(P, T).
inline void find_rT_rP(const Mixture& mx, double rho, double P, const Frac& b, const MixOpt& o, double* rhos,
                       double& T, double& TdR, double& TdP) {
    const double v_mix = 1.0 / rho;
    T = std::isnan(o.T0) ? 300.0 : o.T0;
    init_densities(mx, b, o, rhos);
    DPDT v{NAN, NAN, NAN};
    int counter = 0;
    while (counter < kMaxIter) {
        double err = 0.0;
        v = {0.0, 0.0, 0.0};
        for (int i = 0; i < mx.n; ++i) {
            if (b.has(i)) {
                DPDT vi = sg_volume_PT(mx.m[i], P, T);
                v.val += b.a[i] * vi.val;
                v.dP += b.a[i] * vi.dP;
                v.dT += b.a[i] * vi.dT;
                err = smax(err, std::abs(1.0 - rhos[i] * vi.val));
                rhos[i] = 1.0 / vi.val;
            }
        }
        err = smax(err, std::abs(1.0 - rho * v.val));
        if (err < kEps) break;
        T += (v_mix - v.val) / v.dT;
        ++counter;
    }
    TdR = NAN;
    TdP = NAN;
    if (o.deriv) {
        TdR = -(v.val * v.val) / v.dT;
        TdP = -v.dP / v.dT;
    }
}

// energy_PT of the mixture (mixture_pt.cpp:134-150).
inline DPDT mix_energy_PT(const Mixture& mx, double P, double T, const Frac& b) {
    DPDT e{0.0, 0.0, 0.0};
    for (int i = 0; i < mx.n; ++i)
        if (b.has(i)) {
            DPDT ei = sg_energy_PT(mx.m[i], P, T);
            e.val += b.a[i] * ei.val;
            e.dP += b.a[i] * ei.dP;
            e.dT += b.a[i] * ei.dT;
        }
    return e;
}

// find_reT_rP_old (mixture_pt.cpp:434-446).
inline void find_reT_rP(const Mixture& mx, double rho, double P, const Frac& b, const MixOpt& o, double* rhos,
                        double& e, double& edR, double& edP, double& T) {
    double TdR, TdP;
    find_rT_rP(mx, rho, P, b, o, rhos, T, TdR, TdP);
    DPDT e1 = mix_energy_PT(mx, P, T, b);
    e = e1.val;
    edR = NAN;
    edP = NAN;
    if (o.deriv) {
        edR = e1.dT * TdR;
        edP = e1.dT * TdP + e1.dP;
    }
}

// find_rPT_old (mixture_pt.cpp:448-533): Newton on (P, T) for
// v_mix = sum beta_i v_i(P, T), e_mix = sum beta_i e_i(P, T).
inline void find_rPT(const Mixture& mx, double rho, double e_mix, const Frac& b, const MixOpt& o, double* rhos,
                     double& P, double& T, DPDT* v_last = nullptr, DPDT* e_last = nullptr) {
    const double v_mix = 1.0 / rho;
    P = std::isnan(o.P0) ? 1.0e5 : o.P0;
    T = std::isnan(o.T0) ? 300.0 : o.T0;
    const double P_min = min_pressure(mx, b);
    if (P < P_min) P = P_min + 1;
    init_densities(mx, b, o, rhos);
    int counter = 0;
    DPDT v{NAN, NAN, NAN}, e{NAN, NAN, NAN};
    while (counter < kMaxIter) {
        double err = 0.0;
        v = {0.0, 0.0, 0.0};
        e = {0.0, 0.0, 0.0};
        for (int i = 0; i < mx.n; ++i) {
            if (b.has(i)) {
                DPDT vi = sg_volume_PT(mx.m[i], P, T);
                v.val += b.a[i] * vi.val;
                v.dP += b.a[i] * vi.dP;
                v.dT += b.a[i] * vi.dT;
                err = smax(err, std::abs(1.0 - rhos[i] * vi.val));
                rhos[i] = 1.0 / vi.val;
                DPDT ei = sg_energy_PT(mx.m[i], P, T);
                e.val += b.a[i] * ei.val;
                e.dP += b.a[i] * ei.dP;
                e.dT += b.a[i] * ei.dT;
            }
        }
        err = smax(err, std::abs(1.0 - rho * v.val));
        const double err_e = std::abs(1.0 - e.val / e_mix);
        if (err < kEps && err_e < kEps) break;
        const double inv_D = 1.0 / (v.dT * e.dP - v.dP * e.dT);   // inv_J (mixture_pt.cpp:22-24)
        const double dP = inv_D * ((e_mix - e.val) * v.dT - (v_mix - v.val) * e.dT);
        const double dT = inv_D * ((v_mix - v.val) * e.dP - (e_mix - e.val) * v.dP);
        double dP_lim = dP;
        if (P + dP < P_min) {
            const double a = 0.1;
            dP_lim = (1.0 - a) * P_min + (a - 1.0) * P;
        }
        const double dT_lim = dP != 0.0 ? (dP_lim / dP) * dT : dT;
        P += dP_lim;
        T += dT_lim;
        ++counter;
    }
    if (v_last) *v_last = v;
    if (e_last) *e_last = e;
}

// get_rPT (mixture_pt.cpp:199-212): pure material closed form, else find_rPT.
inline void get_rPT(const Mixture& mx, double rho, double e, const Frac& b, const MixOpt& o, double* rhos,
                    double& P, double& T) {
    const int idx = b.index();
    if (idx >= 0) {
        for (int i = 0; i < kMaxMat; ++i) rhos[i] = NAN;
        rhos[idx] = rho;
        P = sg_pressure_re(mx.m[idx], rho, e);
        T = sg_temperature_rP(mx.m[idx], rho, P);
        return;
    }
    find_rPT(mx, rho, e, b, o, rhos, P, T);
}

// get_reT (mixture_pt.cpp:214-227).
inline void get_reT(const Mixture& mx, double rho, double P, const Frac& b, const MixOpt& o, double* rhos,
                    double& e, double& T) {
    const int idx = b.index();
    if (idx >= 0) {
        for (int i = 0; i < kMaxMat; ++i) rhos[i] = NAN;
        rhos[idx] = rho;
        e = sg_energy_rP(mx.m[idx], rho, P);
        T = sg_temperature_rP(mx.m[idx], rho, P);
        return;
    }
    double edR, edP;
    find_reT_rP(mx, rho, P, b, o, rhos, e, edR, edP, T);
}

// sound_speed_rP (mixture_pt.cpp:101-114).
inline double mix_sound_speed_rP(const Mixture& mx, double rho, double P, const Frac& b, const MixOpt& o) {
    const int idx = b.index();
    if (idx >= 0) return sg_sound_speed_rP(mx.m[idx], rho, P);
    MixOpt od = o;
    od.deriv = true;
    double rhos[kMaxMat], e, edR, edP, T;
    find_reT_rP(mx, rho, P, b, od, rhos, e, edR, edP, T);
    const double c2 = (P / (rho * rho) - edR) / edP;
    return std::sqrt(c2);
}

// mmf::PState (math/cfd/models.h:164-283): 2D, velocity z = 0.
struct MState {
    double rho;
    Vec3 v;
    double P, e, T;
    Frac beta;
    double rhos[kMaxMat];
};

inline double m_E(const MState& z) { return z.e + 0.5 * sqnorm(z.v); }   // models.h:237

// PState::alpha (models.cpp:224-227).
inline double m_alpha(const MState& z, int i) {
    if (std::isnan(z.rhos[i])) return 0.0;
    const double a = z.beta.a[i] * z.rho / z.rhos[i];
    const double lo = (1.0 < a) ? 1.0 : a;   // std::min(a, 1.0)
    return (0.0 < lo) ? lo : 0.0;            // std::max(0.0, lo)
}

// PState::volume_fractions (models.cpp:229-239): unclamped, no normalisation.
inline Frac m_volume_fractions(const MState& z) {
    Frac a;
    for (int i = 0; i < kMaxMat; ++i)
        if (z.beta.has(i)) a.a[i] = z.beta.a[i] * z.rho / z.rhos[i];
    return a;
}

// mmf::QState / mmf::Flux as 9 doubles {rho, m.x, m.y, m.z, E, mass_frac[4]}.
struct MArr { double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; };

inline MArr m_qstate(const MState& z) {   // QState(const PState&) (models.cpp:413-420)
    MArr q;
    q.a[0] = z.rho;
    q.a[1] = z.rho * z.v.x;
    q.a[2] = z.rho * z.v.y;
    q.a[3] = z.rho * z.v.z;
    q.a[4] = z.rho * m_E(z);
    for (int i = 0; i < kMaxMat; ++i) q.a[5 + i] = z.rho * z.beta.a[i];
    return q;
}

inline MArr m_flux(const MState& z) {   // Flux(const PState&) (models.cpp:459-469)
    MArr f;
    f.a[0] = z.rho * z.v.x;
    f.a[1] = z.rho * z.v.x * z.v.x + z.P;
    f.a[2] = z.rho * z.v.x * z.v.y;
    f.a[3] = z.rho * z.v.x * z.v.z;
    f.a[4] = (z.rho * m_E(z) + z.P) * z.v.x;
    for (int i = 0; i < kMaxMat; ++i) f.a[5 + i] = z.rho * z.v.x * z.beta.a[i];
    return f;
}

// PState(const QState&, mixture, P0, T0, rhos) (models.cpp:187-201).
inline MState m_from_q(const MArr& q, const Mixture& mx, double P0, double T0, const double* rhos0) {
    MState z;
    z.rho = q.a[0];
    z.v = {q.a[1] / z.rho, q.a[2] / z.rho, q.a[3] / z.rho};
    z.e = q.a[4] / z.rho - 0.5 * sqnorm(z.v);
    for (int i = 0; i < mx.n; ++i) z.beta.a[i] = q.a[5 + i] / z.rho;
    z.beta.cutoff(1.0e-12);
    z.beta.normalize();
    MixOpt o;
    o.P0 = P0;
    o.T0 = T0;
    o.rhos = rhos0;
    get_rPT(mx, z.rho, z.e, z.beta, o, z.rhos, z.P, z.T);
    return z;
}

// PState(rho, v, P, beta, mixture) (models.cpp:176-185).
inline MState m_from_rvp(const Mixture& mx, double rho, double u, double v, double P, const Frac& beta) {
    MState z;
    z.rho = rho;
    z.v = {u, v, 0.0};
    z.P = P;
    z.beta = beta;
    MixOpt o;
    get_reT(mx, rho, P, beta, o, z.rhos, z.e, z.T);
    return z;
}

// HLLC::calc_flux, mmf (math/cfd/flux/hllc.cpp:339-421); states in the face frame.
inline MArr m_hllc(const MState& zL, const MState& zR, const Mixture& mx) {
    MixOpt oL, oR;
    oL.T0 = zL.T;
    oL.rhos = zL.rhos;
    oR.T0 = zR.T;
    oR.rhos = zR.rhos;
    const double c_L = mix_sound_speed_rP(mx, zL.rho, zL.P, zL.beta, oL);
    const double c_R = mix_sound_speed_rP(mx, zR.rho, zR.P, zR.beta, oR);
    double S_L = smin(zL.v.x - c_L, zR.v.x - c_L);
    double S_R = smax(zL.v.x + c_R, zR.v.x + c_R);
    if (S_L >= S_R || zL.beta.index() == zR.beta.index()) {
        S_L = smin(zL.v.x - c_L, zR.v.x - c_R);
        S_R = smax(zL.v.x + c_L, zR.v.x + c_R);
    }
    if (S_L >= 0.0) return m_flux(zL);
    if (S_R <= 0.0) return m_flux(zR);
    const double a_L = zL.rho * (S_L - zL.v.x);
    const double a_R = zR.rho * (S_R - zR.v.x);
    double S_C = (zL.P - zR.P + a_R * zR.v.x - a_L * zL.v.x) / (a_R - a_L);
    if (S_C < S_L) S_C = S_L + 1.0e-9 * smax(c_L, c_R);
    if (S_C > S_R) S_C = S_R - 1.0e-9 * smax(c_L, c_R);
    const double rho_sL = a_L / (S_L - S_C);
    const double rho_sR = a_R / (S_R - S_C);
    const double E_sL = m_E(zL) + (S_C - zL.v.x) * (S_C + zL.P / a_L);
    const double E_sR = m_E(zR) + (S_C - zR.v.x) * (S_C + zR.P / a_R);
    MArr QsL, QsR;
    QsL.a[0] = rho_sL;
    QsR.a[0] = rho_sR;
    QsL.a[1] = rho_sL * S_C; QsL.a[2] = rho_sL * zL.v.y; QsL.a[3] = rho_sL * zL.v.z;
    QsR.a[1] = rho_sR * S_C; QsR.a[2] = rho_sR * zR.v.y; QsR.a[3] = rho_sR * zR.v.z;
    QsL.a[4] = rho_sL * E_sL;
    QsR.a[4] = rho_sR * E_sR;
    for (int i = 0; i < mx.n; ++i) {
        QsL.a[5 + i] = rho_sL * zL.beta.a[i];
        QsR.a[5 + i] = rho_sR * zR.beta.a[i];
    }
    const MArr QL = m_qstate(zL), QR = m_qstate(zR), FL = m_flux(zL), FR = m_flux(zR);
    const double aSC = std::abs(S_C);
    MArr F;
    for (int k = 0; k < 9; ++k) {
        double t = FL.a[k] + FR.a[k];
        t = t + S_L * (QsL.a[k] - QL.a[k]);
        t = t + aSC * (QsL.a[k] - QsR.a[k]);
        t = t + S_R * (QsR.a[k] - QR.a[k]);
        F.a[k] = 0.5 * t;
        if (std::isnan(F.a[k])) g_numeric_error = 1;   // hllc.cpp:407-417 throws
    }
    return F;
}

// Rusanov::calc_flux, mmf (math/cfd/flux/rusanov.cpp:46-83).
inline MArr m_rusanov(const MState& zL, const MState& zR, const Mixture& mx) {
    MixOpt oL, oR;
    oL.T0 = zL.T;
    oL.rhos = zL.rhos;
    oR.T0 = zR.T;
    oR.rhos = zR.rhos;
    const double c_L = mix_sound_speed_rP(mx, zL.rho, zL.P, zL.beta, oL);
    const double c_R = mix_sound_speed_rP(mx, zR.rho, zR.P, zR.beta, oR);
    if (std::isnan(c_L) || std::isnan(c_R)) g_numeric_error = 1;   // :53-59 throws
    const double S_L = smin(smin(zL.v.x - c_L, zR.v.x - c_R), 0.0);   // std::min({...})
    const double S_R = smax(smax(zL.v.x + c_L, zR.v.x + c_R), 0.0);
    const double S = smax(std::abs(S_L), S_R);
    const MArr QL = m_qstate(zL), QR = m_qstate(zR), FL = m_flux(zL), FR = m_flux(zR);
    MArr F;
    for (int k = 0; k < 9; ++k) {
        F.a[k] = 0.5 * (FL.a[k] + FR.a[k]) + 0.5 * S * (QL.a[k] - QR.a[k]);
        if (k < 5 && !std::isfinite(F.a[k])) g_numeric_error = 1;   // Flux::is_bad (models.cpp:535-541)
    }
    return F;
}

// HLL::calc_flux, mmf (math/cfd/flux/hll.cpp:125-160).
inline MArr m_hll(const MState& zL, const MState& zR, const Mixture& mx) {
    MixOpt oL, oR;
    oL.T0 = zL.T;
    oL.rhos = zL.rhos;
    oR.T0 = zR.T;
    oR.rhos = zR.rhos;
    const double c_L = mix_sound_speed_rP(mx, zL.rho, zL.P, zL.beta, oL);
    const double c_R = mix_sound_speed_rP(mx, zR.rho, zR.P, zR.beta, oR);
    double S_L = smin(zL.v.x - c_L, zR.v.x - c_L);
    double S_R = smax(zL.v.x + c_R, zR.v.x + c_R);
    if (S_L >= S_R) {
        S_L = smin(zL.v.x - c_L, zR.v.x - c_R);
        S_R = smax(zL.v.x + c_L, zR.v.x + c_R);
    }
    const MArr QL = m_qstate(zL), QR = m_qstate(zR), FL = m_flux(zL), FR = m_flux(zR);
    MArr F;
    for (int k = 0; k < 9; ++k) {
        F.a[k] = (S_R * FL.a[k] - S_L * FR.a[k] + S_L * S_R * (QR.a[k] - QL.a[k])) / (S_R - S_L);
        if (std::isnan(F.a[k])) g_numeric_error = 1;   // :151-158 throws
    }
    return F;
}

// boundary_value, mmf (math/solver/mm_fluid.cpp:98-108).
inline MState m_boundary_value(const MState& zc, const Vec3& n, int flag) {
    if (flag != ORC_WALL) return zc;
    MState zn = zc;
    const Vec3 Vn = scale(dot(zc.v, n), n);
    zn.v = {zc.v.x - 2.0 * Vn.x, zc.v.y - 2.0 * Vn.y, zc.v.z - 2.0 * Vn.z};
    return zn;
}

inline MState m_in_local(const MState& z, const Vec3& n) {
    MState r = z;
    r.v = rot_local(z.v, n);
    return r;
}

// ---- PLIC formulas (geom/sections.cpp) and math/funcs.h helpers ----------
inline double f_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }   // funcs.h:21-23
inline double f_heav(double x) { return x > 0.0 ? 1.0 : 0.0; }                        // funcs.h:27-29
inline void f_sorted(double a, double b, double& lo, double& hi) {                    // funcs.h:32-35
    if (a < b) { lo = a; hi = b; }
    else { lo = b; hi = a; }
}
inline double f_between(double x, double x1, double x2) {   // funcs.h:47-50
    double lo, hi;
    f_sorted(x1, x2, lo, hi);
    return smax(lo, smin(x, hi));
}

double quad_volume_fraction(double p, double nx, double ny, double A, double B) {   // sections.cpp:13-23
    double a, b;
    f_sorted(A * std::abs(nx), B * std::abs(ny), a, b);
    if (std::abs(p) <= 0.5 * (b - a)) return 0.5 + p / b;
    if (std::abs(p) < 0.5 * (b + a)) {
        const double t = std::abs(p) - 0.5 * (a + b);
        return f_heav(p) - 0.5 * f_sign(p) * (t * t) / (a * b);
    }
    return f_heav(p);
}

double quad_find_section(double alpha, double nx, double ny, double A, double B) {   // sections.cpp:25-37
    double a, b;
    f_sorted(A * std::abs(nx), B * std::abs(ny), a, b);
    if (alpha <= 0.5 * a / b) return -0.5 * (a + b) + std::sqrt(2.0 * alpha * a * b);
    if (alpha < 1.0 - 0.5 * a / b) return (alpha - 0.5) * b;
    return +0.5 * (a + b) - std::sqrt(2.0 * (1.0 - alpha) * a * b);
}

inline void minmax_unit(double a1, double a2, double& lo, double& hi) {   // sections.cpp:42-48
    if (a1 < a2) { lo = smax(0.0, a1); hi = smin(a2, 1.0); }
    else { lo = smax(0.0, a2); hi = smin(a1, 1.0); }
}

double face_fraction_v3(double a1, double a2) {   // sections.cpp:52-62
    double lo, hi;
    minmax_unit(a1, a2, lo, hi);
    if (lo == 0.0) return 0.0;
    if (hi == 1.0) return 1.0;
    if (a1 + a2 < 1.0) return hi;
    if (a1 + a2 > 1.0) return lo;
    return 0.5 * (a1 + a2);
}

double face_fraction_v5(double a1, double a2) {   // sections.cpp:64-71
    double lo, hi;
    minmax_unit(a1, a2, lo, hi);
    if (lo == 0.0) return 0.0;
    if (hi == 1.0) return 1.0;
    return 0.5 * (a1 + a2);
}

double face_fraction_s(double a1, double a2) {   // sections.cpp:73-84
    double lo, hi;
    minmax_unit(a1, a2, lo, hi);
    double a_sig = lo / (1.0 - (hi - lo));
    if (std::isnan(a_sig)) a_sig = 0.5;
    return f_between(a_sig, lo, hi);
}

double face_fraction(double a1, double a2) {   // sections.cpp:86-104
    double lo, hi;
    minmax_unit(a1, a2, lo, hi);
    if (3.0 * lo >= hi && lo >= 3.0 * hi - 2.0) return 0.5 * (lo + hi);
    if (hi <= 0.5 || lo * (1.0 - hi) >= (hi - 0.5) * (hi - 0.5))
        return -2.0 * lo + 2.0 * std::sqrt(lo * (lo + hi));
    if (lo >= 0.5 || lo * (1.0 - hi) >= (lo - 0.5) * (lo - 0.5))
        return 3.0 - 2.0 * hi - 2.0 * std::sqrt((1.0 - hi) * (2.0 - lo - hi));
    if (lo == 0.0 && hi == 1.0) return 0.5;
    return std::sqrt(lo) / (std::sqrt(1.0 - hi) + std::sqrt(lo));
}

double average_flux(double alpha, double cs, double CFL) {   // sections.cpp:125-148
    double a1, b1, a2, b2;
    f_sorted(std::abs(cs), std::sqrt(1.0 - cs * cs), a1, b1);
    f_sorted(CFL * std::abs(cs), std::sqrt(1.0 - cs * cs), a2, b2);
    double P1;
    if (std::abs(2 * alpha - 1) <= 1.0 - a1 / b1) P1 = (alpha - 0.5) * b1;
    else P1 = f_sign(alpha - 0.5) * (0.5 * (a1 + b1) - std::sqrt((1.0 - std::abs(2.0 * alpha - 1.0)) * a1 * b1));
    const double P2 = P1 + 0.5 * (CFL - 1.0) * cs;
    if (std::abs(P2) <= 0.5 * (b2 - a2)) return 0.5 + P2 / b2;
    if (std::abs(P2) < 0.5 * (b2 + a2)) {
        const double t = std::abs(P2) - 0.5 * (a2 + b2);
        return f_heav(P2) - 0.5 * f_sign(P2) * (t * t) / (a2 * b2);
    }
    return f_heav(P2);
}

// sound_speed_re of the mixture (mixture_pt.cpp:86-99): pure material
// sqrt(g (g - 1) ((e - e0) - P0 / rho)) (stiffened_gas.cpp:144-146), else
// find_rPT with derivatives (P.dE = inv_D v.dT, P.dR = inv_D e.dT v^2,
// mixture_pt.cpp:525-529) and c^2 = P.dR + P P.dE / rho^2.
inline double mix_sound_speed_re(const Mixture& mx, double rho, double e, const Frac& b) {
    const int idx = b.index();
    if (idx >= 0) {
        const SgMat& m = mx.m[idx];
        return std::sqrt(m.g * (m.g - 1.0) * ((e - m.e0) - m.P0 / rho));
    }
    MixOpt o;
    o.deriv = true;
    double rhos[kMaxMat], P, T;
    DPDT v, ee;
    find_rPT(mx, rho, e, b, o, rhos, P, T, &v, &ee);
    const double inv_D = 1.0 / (v.dT * ee.dP - v.dP * ee.dT);
    const double PdE = inv_D * v.dT;
    const double PdR = inv_D * ee.dT * (v.val * v.val);
    const double c2 = PdR + P * PdE / (rho * rho);
    return std::sqrt(c2);
}

// Planted mistakes for the tests that show the composite-flux checks can
// fail (0 = none, the restatement as written): 1 swaps the inverse problem
// of the aL == 0 branch of calc_crp_flux for a classic one, 2 flips the
// mix-to-clear / clear-to-mix test of crp_classic, 3 gives the mixture
// part of PState::split the energy of the first of its materials alone,
// 4 takes the right star state where the first branch of crp_classic
// takes the left one.
inline int g_orc_plant = 0;

// mmf::WaveConfig3 (hllc.h:21-27).
struct Wave {
    double SL, SC, SR;
    MArr QsL, FsL, QsR, FsR;
};

// HLLC::wave_config from conserved vectors and fluxes (hllc.cpp:256-335).
inline Wave wave_config(const Mixture& mx, const MArr& QL, const MArr& FL, const MArr& QR, const MArr& FR) {
    const Vec3 vL{QL.a[1] / QL.a[0], QL.a[2] / QL.a[0], QL.a[3] / QL.a[0]};
    const Vec3 vR{QR.a[1] / QR.a[0], QR.a[2] / QR.a[0], QR.a[3] / QR.a[0]};
    const double eL = QL.a[4] / QL.a[0] - 0.5 * sqnorm(vL);
    const double eR = QR.a[4] / QR.a[0] - 0.5 * sqnorm(vR);
    Frac bL, bR;
    for (int i = 0; i < mx.n; ++i) {
        bL.a[i] = QL.a[5 + i] / QL.a[0];
        bR.a[i] = QR.a[5 + i] / QR.a[0];
    }
    const double cL = mix_sound_speed_re(mx, QL.a[0], eL, bL);
    const double cR = mix_sound_speed_re(mx, QR.a[0], eR, bR);
    double SL = smin(vL.x - cL, vR.x - cL);
    double SR = smax(vL.x + cR, vR.x + cR);
    if (SL >= SR || bL.index() == bR.index()) {
        SL = smin(vL.x - cL, vR.x - cR);
        SR = smax(vL.x + cL, vR.x + cR);
    }
    MArr aL, aR;
    for (int k = 0; k < 9; ++k) {
        aL.a[k] = FL.a[k] - SL * QL.a[k];
        aR.a[k] = FR.a[k] - SR * QR.a[k];
    }
    double SC = (aR.a[1] - aL.a[1]) / (aR.a[0] - aL.a[0]);
    if (SC < SL) SC = SL + 1.0e-9 * smax(cL, cR);
    if (SC > SR) SC = SR - 1.0e-9 * smax(cL, cR);
    const double rsL = aL.a[0] / (SC - SL);
    const double rsR = aR.a[0] / (SC - SR);
    const double EsL = SC * SC + (aL.a[4] - SC * aL.a[1]) / aL.a[0];
    const double EsR = SC * SC + (aR.a[4] - SC * aR.a[1]) / aR.a[0];
    Wave w;
    w.SL = SL;
    w.SC = SC;
    w.SR = SR;
    w.QsL.a[0] = rsL;
    w.QsR.a[0] = rsR;
    w.QsL.a[1] = rsL * SC; w.QsL.a[2] = rsL * vL.y; w.QsL.a[3] = rsL * vL.z;
    w.QsR.a[1] = rsR * SC; w.QsR.a[2] = rsR * vR.y; w.QsR.a[3] = rsR * vR.z;
    w.QsL.a[4] = rsL * EsL;
    w.QsR.a[4] = rsR * EsR;
    for (int i = 0; i < mx.n; ++i) {
        w.QsL.a[5 + i] = rsL * bL.a[i];
        w.QsR.a[5 + i] = rsR * bR.a[i];
    }
    for (int k = 0; k < 9; ++k) {
        w.FsL.a[k] = aL.a[k] + SL * w.QsL.a[k];
        w.FsR.a[k] = aR.a[k] + SR * w.QsR.a[k];
    }
    return w;
}

// Characteristics in the (x, t) plane (crp_flux.cpp:13-55).
struct XT { double x, t; };
struct Chr {
    double x, t, S;
    double edge_t() const {
        const double ts = (S * t - x) / S;
        return ts < t ? std::numeric_limits<double>::infinity() : ts;
    }
    XT cross(const Chr& c) const {
        constexpr double inf = std::numeric_limits<double>::infinity();
        if (S == c.S) return {0.5 * (x + c.x), inf};
        XT p;
        p.t = (c.x - x + S * t - c.S * c.t) / (S - c.S);
        if (p.t <= smax(t, c.t)) return {0.5 * (x + c.x), inf};
        p.x = 0.5 * (x + c.x + S * (p.t - t) + c.S * (p.t - c.t));
        return p;
    }
};

// F (t1 - t0) - Q (x1 - x0) for the segment from point 0 to point 1.
inline MArr seg(const MArr& F, const MArr& Q, double dtv, double dxv) {
    MArr r;
    for (int k = 0; k < 9; ++k) r.a[k] = F.a[k] * dtv - Q.a[k] * dxv;
    return r;
}

// CrpFlux::classic, ZPP version (crp_flux.cpp:234-489): a pure material A
// at distance delta behind the face inside the left cell's mixture B, and
// the right state RB; the face flux averaged over dt from the wave pattern.
// err: set for the reference's "very strange case" (it throws).
inline MArr crp_classic(const MState& zLA, const MState& zLB, const MState& zRB, const Mixture& mx, double delta,
                        double dt, int& err) {
    const MArr qLB = m_qstate(zLB), fLB = m_flux(zLB), Q_R = m_qstate(zRB), F_R = m_flux(zRB);
    const Wave w0 = wave_config(mx, qLB, fLB, Q_R, F_R);
    const Chr C_0{-delta, 0.0, zLA.v.x};
    const Chr C_0L{0.0, 0.0, w0.SL};
    const XT O1 = C_0.cross(C_0L);
    const XT O0{0.0, 0.0}, Odt{0.0, dt};
    auto total2 = [&](const MArr& G1, const MArr& G2) {
        MArr r;
        for (int k = 0; k < 9; ++k) r.a[k] = (G1.a[k] + G2.a[k]) / (0.0 - dt);
        return r;
    };
    if (O1.x >= 0.0) {
        const double t2u = C_0.edge_t();
        if (dt <= t2u) return fLB;
        const MArr qLA = m_qstate(zLA), fLA = m_flux(zLA);
        const Wave w1 = wave_config(mx, qLA, fLA, w0.QsL, w0.FsL);
        const double lb = -O1.x / (dt - O1.t);
        MArr F = fLA, Q = qLA;
        if (dt > O1.t) {
            if (lb < w1.SL) { F = fLA; Q = qLA; }
            else if (lb < w1.SC) {
                F = g_orc_plant == 4 ? w1.FsR : w1.FsL;
                Q = g_orc_plant == 4 ? w1.QsR : w1.QsL;
            }
            else if (lb < w1.SR) { F = w1.FsR; Q = w1.QsR; }
            else { F = w0.FsL; Q = w0.QsL; }
        }
        const MArr G_OdtO1 = seg(F, Q, O1.t - Odt.t, O1.x - Odt.x);
        const MArr G_O1O0 = seg(fLB, qLB, O0.t - O1.t, O0.x - O1.x);
        return total2(G_OdtO1, G_O1O0);
    }
    if ((w0.SC >= 0.0) != (g_orc_plant == 2)) {   // mix to clear
        if (dt <= O1.t && O1.x < 0.0) return w0.FsL;
        const MArr Q_L = m_qstate(zLA), F_L = m_flux(zLA);
        const Wave w1 = wave_config(mx, Q_L, F_L, w0.QsL, w0.FsL);
        const Chr C_1R{O1.x, O1.t, w1.SR};
        const double t1 = C_1R.edge_t();
        if (dt <= t1 && O1.x < 0.0) return w0.FsL;
        const double lb = -O1.x / (dt - O1.t);
        const MArr& F = lb < w1.SL ? F_L : (lb < w1.SC ? w1.FsL : w1.FsR);
        const MArr& Q = lb < w1.SL ? Q_L : (lb < w1.SC ? w1.QsL : w1.QsR);
        const MArr G_OdtO1 = seg(F, Q, O1.t - Odt.t, O1.x - Odt.x);
        const MArr G_O1O = seg(w0.FsL, w0.QsL, O0.t - O1.t, O0.x - O1.x);
        return total2(G_OdtO1, G_O1O);
    }
    // clear to mix
    if (O1.t >= dt) return w0.FsR;
    const MArr Q_L = m_qstate(zLA), F_L = m_flux(zLA);
    const Wave w1 = wave_config(mx, Q_L, F_L, w0.QsL, w0.FsL);
    const Chr C_0C{0.0, 0.0, w0.SC};
    const Chr C_1R{O1.x, O1.t, w1.SR};
    const XT O2 = C_1R.cross(C_0C);
    if (O2.t > dt || O2.x > 0.0) return w0.FsR;
    const Wave w2 = wave_config(mx, w1.QsR, w1.FsR, w0.QsR, w0.FsR);
    const Chr C_2L{O2.x, O2.t, w2.SL};
    const Chr C_1C{O1.x, O1.t, w1.SC};
    const XT O3 = C_1C.cross(C_2L);
    if (O3.x >= 0.0) {
        err = 1;   // crp_flux.cpp:419-421 throws
        return w0.FsR;
    }
    const Wave w3 = wave_config(mx, w1.QsL, w1.FsL, w2.QsL, w2.FsL);
    const Chr C_2R{O2.x, O2.t, w2.SR};
    const Chr C_3R{O3.x, O3.t, w3.SR};
    const double t1 = C_2R.edge_t();
    const double t4 = C_3R.edge_t();
    if (dt <= t4) {
        if (dt <= t1) return w0.FsR;
        const double lb = -O2.x / (dt - O2.t);
        const MArr& F = lb < w2.SC ? w2.FsL : (lb < w2.SR ? w2.FsR : w0.FsR);
        const MArr& Q = lb < w2.SC ? w2.QsL : (lb < w2.SR ? w2.QsR : w0.QsR);
        const MArr G_OdtO2 = seg(F, Q, O2.t - Odt.t, O2.x - Odt.x);
        const MArr G_O2O0 = seg(w0.FsR, w0.QsR, O0.t - O2.t, O0.x - O2.x);
        return total2(G_OdtO2, G_O2O0);
    }
    const double lb = -O3.x / (dt - O3.t);
    const MArr& F = lb < w3.SL ? w1.FsL : (lb < w3.SC ? w3.FsL : w3.FsR);
    const MArr& Q = lb < w3.SL ? w1.QsL : (lb < w3.SC ? w3.QsL : w3.QsR);
    const MArr G_OdtO3 = seg(F, Q, O3.t - Odt.t, O3.x - Odt.x);
    const MArr G_O3O2 = seg(w2.FsL, w2.QsL, O2.t - O3.t, O2.x - O3.x);
    const MArr G_O2O0 = seg(w0.FsR, w0.QsR, O0.t - O2.t, O0.x - O2.x);
    MArr r;
    for (int k = 0; k < 9; ++k) r.a[k] = (G_OdtO3.a[k] + G_O3O2.a[k] + G_O2O0.a[k]) / (0.0 - dt);
    return r;
}

// PState::inverse / Flux::inverse (models.cpp:219-221, 506-515).
inline MState m_inverse(MState z) {
    z.v.x = -z.v.x;
    return z;
}
inline MArr f_inverse(MArr f) {
    f.a[0] = -f.a[0];
    f.a[2] = -f.a[2];
    f.a[3] = -f.a[3];
    f.a[4] = -f.a[4];
    for (int i = 0; i < kMaxMat; ++i) f.a[5 + i] = -f.a[5 + i];
    return f;
}

// CrpFlux::inverse (crp_flux.cpp:493-512): the mirrored classic problem.
inline MArr crp_inverse(const MState& zLA, const MState& zRA, const MState& zRB, const Mixture& mx, double delta,
                        double dt, int& err) {
    return f_inverse(crp_classic(m_inverse(zRB), m_inverse(zRA), m_inverse(zLA), mx, delta, dt, err));
}

// PState::split, ZPP version (models.cpp:261-335): pure material iA and the
// mixture of the others at the same velocity, pressure and temperature.
inline void m_split(const MState& z, const Mixture& mx, int iA, MState& zA, MState& zB) {
    const double energy_A = sg_energy_rT(mx.m[iA], z.rhos[iA], z.T);
    zA.rho = z.rhos[iA];
    zA.v = z.v;
    zA.P = z.P;
    zA.e = energy_A;
    zA.T = z.T;
    zA.beta = Frac();
    zA.beta.a[iA] = 1.0;
    for (int i = 0; i < kMaxMat; ++i) zA.rhos[i] = NAN;
    zA.rhos[iA] = z.rhos[iA];

    zB.rho = NAN;
    zB.v = z.v;
    zB.P = z.P;
    zB.e = NAN;
    zB.T = z.T;
    zB.beta = Frac();
    for (int i = 0; i < kMaxMat; ++i) zB.rhos[i] = NAN;

    const double beta_B = 1.0 - z.beta.a[iA];
    Frac alpha;
    for (int i = 0; i < kMaxMat; ++i) {
        if (z.beta.has(i) && i != iA) {
            zB.rhos[i] = z.rhos[i];
            zB.beta.a[i] = z.beta.a[i] / beta_B;
        }
        if (z.beta.has(i)) alpha.a[i] = m_alpha(z, i);
    }
    zB.beta.normalize();
    const double alpha_B = 1.0 - alpha.a[iA];
    alpha.a[iA] = 0.0;
    for (int i = 0; i < kMaxMat; ++i)
        if (z.beta.has(i) && i != iA) alpha.a[i] = alpha.a[i] / alpha_B;
    alpha.normalize();
    zB.rho = 0.0;
    for (int i = 0; i < kMaxMat; ++i)
        if (z.beta.has(i) && i != iA) zB.rho = zB.rho + z.rhos[i] * alpha.a[i];
    zB.e = 0.0;
    for (int i = 0; i < kMaxMat; ++i)
        if (z.beta.has(i) && i != iA) {
            zB.beta.a[i] = alpha.a[i] * z.rhos[i] / zB.rho;
            zB.e += zB.beta.a[i] * sg_energy_rT(mx.m[i], z.rhos[i], z.T);
        }
    if (g_orc_plant == 3)
        for (int i = 0; i < kMaxMat; ++i)
            if (z.beta.has(i) && i != iA) {
                zB.e = sg_energy_rT(mx.m[i], z.rhos[i], z.T);
                break;
            }
}

// master_component (mm_fluid.cpp:201-226): largest volume fraction of the upwind side.
inline int master_component(const MState& zm, const MState& zp) {
    const double vi = 0.5 * (zm.v.x + zp.v.x);
    const MState& z = vi > 0.0 ? zm : zp;
    int master = 0;
    double alpha_max = 0.0;
    for (int i = 0; i < kMaxMat; ++i)
        if (z.beta.has(i)) {
            const double a = m_alpha(z, i);
            if (a > alpha_max) {
                alpha_max = a;
                master = i;
            }
        }
    return master;
}

// MmFluid::calc_crp_flux (mm_fluid.cpp:259-337).
inline MArr calc_crp_flux(const MState& zL, const MState& zR, double hL, double hR, int iA, double a_sig, double dt,
                          const Mixture& mx, int& err) {
    const double aL = m_alpha(zL, iA), aR = m_alpha(zR, iA);
    if ((aL == 0.0 || aL == 1.0) && (aR == 0.0 || aR == 1.0)) return m_hllc(zL, zR, mx);
    const double delta_L = (aL < a_sig ? aL / a_sig : (1.0 - aL) / (1.0 - a_sig)) * hL;
    const double delta_R = (aR < a_sig ? aR / a_sig : (1.0 - aR) / (1.0 - a_sig)) * hR;
    MArr fA, fB;
    if (aL == 0.0) {
        MState zRA, zRB;
        m_split(zR, mx, iA, zRA, zRB);
        fA = m_hllc(zL, zRA, mx);
        fB = g_orc_plant == 1 ? crp_classic(zL, zRB, zRA, mx, delta_R, dt, err)
                              : crp_inverse(zL, zRB, zRA, mx, delta_R, dt, err);
    } else if (aL == 1.0) {
        MState zRA, zRB;
        m_split(zR, mx, iA, zRA, zRB);
        fA = crp_inverse(zL, zRA, zRB, mx, delta_R, dt, err);
        fB = m_hllc(zL, zRB, mx);
    } else if (aR == 0.0) {
        MState zLA, zLB;
        m_split(zL, mx, iA, zLA, zLB);
        fA = m_hllc(zLA, zR, mx);
        fB = crp_classic(zLA, zLB, zR, mx, delta_L, dt, err);
    } else if (aR == 1.0) {
        MState zLA, zLB;
        m_split(zL, mx, iA, zLA, zLB);
        fA = crp_classic(zLB, zLA, zR, mx, delta_L, dt, err);
        fB = m_hllc(zLB, zR, mx);
    } else {
        MState zLA, zLB, zRA, zRB;
        m_split(zL, mx, iA, zLA, zLB);
        m_split(zR, mx, iA, zRA, zRB);
        if (aL < aR) {
            fA = crp_classic(zLB, zLA, zRA, mx, delta_L, dt, err);
            fB = crp_inverse(zLB, zRB, zRA, mx, delta_R, dt, err);
        } else {
            fA = crp_inverse(zLA, zRA, zRB, mx, delta_R, dt, err);
            fB = crp_classic(zLA, zLB, zRB, mx, delta_L, dt, err);
        }
    }
    MArr r;
    for (int k = 0; k < 9; ++k) r.a[k] = a_sig * fA.a[k] + (1.0 - a_sig) * fB.a[k];
    return r;
}

}  // namespace orcmm

using namespace orcmm;

// =============================================================== MmFluid
// Gradient of the PState components (gradient.h:226-282, 412-474 work
// component by component): rho, v.x, v.y, e, beta[i] (what the split
// distributor reads), then P, T and the true densities rhos[i] (which the
// second-order face extrapolation also carries: FaceExtra::Direct,
// face_extra.h:122-184, moves every component).  v.z is zero throughout,
// and so is its gradient.  A true density is NaN where its material is
// absent, so its gradient is NaN wherever the stencil meets such a cell.
constexpr int kMG = 6 + 2 * kMaxMat;
// CrpMode (mm_fluid.h:36-42) and Fluxes (num_flux.h:11-21) by their
// positions in the reference's enums.
constexpr int kCrpNone = 0, kCrpV3 = 1, kCrpV5 = 2, kCrpVS = 3, kCrpPlic = 4;
constexpr int kFluxGodunov = 0, kFluxRusanov = 1, kFluxHll = 2, kFluxHllc = 3;
struct MG { double a[kMG] = {}; };

struct orc_mm {
    orc_mesh* m;
    Mixture mix;
    int crp = kCrpPlic;   // CrpMode (mm_fluid.h:36-42): the MmFluid default PLIC (mm_fluid.cpp:18)
    int flux = kFluxHllc; // m_nf: HLLC, or HLL / Rusanov with CrpMode::NONE (set_method, mm_fluid.cpp:50-61)
    int split = 0;        // DirSplit (mm_fluid.h:29-33): NONE, SIMPLE, STRANG
    int limiter = ORC_LIM_MC;   // m_limiter = "MC" (mm_fluid.cpp:20)
    int acc = 1;                // m_acc (mm_fluid.h:157), set_accuracy (mm_fluid.cpp:46-48)
    double cfl, dt, max_dt;
    std::vector<MState> init, half, next;   // MmFluid::Parts init / half / next (mm_fluid.h:53-67)
    std::vector<MG> d_dx, d_dy, d_dz; // Parts d_dx / d_dy / d_dz (components of MG)
    std::vector<double> nrm_x, nrm_y, pos;   // interface normals / positions [c][mat]
    int nthreads = 1;   // OpenMP threads of the per-cell loops (each cell writes only itself)
};

namespace {

inline double mg_comp(const MState& z, int k) {
    switch (k) {
        case 0: return z.rho;
        case 1: return z.v.x;
        case 2: return z.v.y;
        case 3: return z.e;
        case 8: return z.P;
        case 9: return z.T;
        default: return k < 8 ? z.beta.a[k - 4] : z.rhos[k - 10];
    }
}

// MmFluid::compute_grad (mm_fluid.cpp:439-466): with a CrpMode other than
// NONE the gradient is zero in a mixed cell and in a cell with a mixed
// face neighbour (a boundary face's neighbour is the cell itself); else
// gradient::LSM (gradient.h:226-282) and gradient::limiting (:412-474) with
// the mm boundary_value (mm_fluid.cpp:98-108), as grad_cell does for SmFluid.
void mm_grad_cell(const orc_mm& S, int c, MG& gx, MG& gy, MG& gz) {
    const orc_mesh& m = *S.m;
    const MState zc = S.init[c];
    gx = MG{};
    gy = MG{};
    gz = MG{};
    if (S.crp != 0) {
        bool zero = zc.beta.index() < 0;
        for (int s8 = 0; s8 < 8 && !zero; ++s8) {
            const size_t f = 8 * (size_t)c + s8;
            const int b = m.bnd[f];
            if (b == ORC_UNDEFINED || slot_is_boundary(b)) continue;
            if (S.init[m.adj[f]].beta.index() < 0) zero = true;
        }
        if (zero) return;   // PState::Zero()
    }
    const double w0 = 1.0;
    const Vec3 cell_c{m.cx[c], m.cy[c], 0.0};
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double Fx[kMG] = {0}, Fy[kMG] = {0}, Fz[kMG] = {0};
    struct FaceGeo { Vec3 dr, drn; double w; MState zn; };
    FaceGeo fg[8];
    int nf = 0;
    for (int s8 = 0; s8 < 8; ++s8) {
        const size_t f = 8 * (size_t)c + s8;
        const int b = m.bnd[f];
        if (b == ORC_UNDEFINED) continue;
        const Vec3 normal{m.fnx[f], m.fny[f], 0.0};
        const MState zn = !slot_is_boundary(b) ? S.init[m.adj[f]] : m_boundary_value(zc, normal, b);
        const Vec3 nc = (b == ORC_INNER) ? Vec3{m.cx[m.adj[f]], m.cy[m.adj[f]], 0.0} : symm_point(m, (int)f, cell_c);
        const Vec3 dr = nc - cell_c;
        const Vec3 drn = scale(dot(dr, normal), normal);
        const double w = w0 * m.area[f] / std::pow(norm(drn), 3);
        const double wd[3] = {w * drn.x, w * drn.y, w * drn.z};
        const double drv[3] = {dr.x, dr.y, dr.z};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] = A[i][j] + wd[i] * drv[j];
        for (int k = 0; k < kMG; ++k) {
            const double dF = mg_comp(zn, k) - mg_comp(zc, k);
            Fx[k] = Fx[k] + (w * dF) * drn.x;
            Fy[k] = Fy[k] + (w * dF) * drn.y;
            Fz[k] = Fz[k] + (w * dF) * drn.z;
        }
        fg[nf++] = {dr, drn, w, zn};
    }
    if (std::abs(A[2][2]) < 1e-8) {
        A[2][2] = 1.0;
        if (std::abs(A[1][1]) < 1e-8) A[1][1] = 1.0;
    }
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1];
    };
    const double c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    const double det = (c00 * A[0][0] + c10 * A[1][0]) + c20 * A[2][0];
    const double invdet = 1.0 / det;
    double B[3][3];
    B[0][0] = c00 * invdet; B[0][1] = c10 * invdet; B[0][2] = c20 * invdet;
    B[1][0] = cof(0, 1) * invdet; B[1][1] = cof(1, 1) * invdet; B[1][2] = cof(2, 1) * invdet;
    B[2][0] = cof(0, 2) * invdet; B[2][1] = cof(1, 2) * invdet; B[2][2] = cof(2, 2) * invdet;
    MG ux, uy, uz;
    for (int k = 0; k < kMG; ++k) {
        ux.a[k] = (B[0][0] * Fx[k] + B[0][1] * Fy[k]) + B[0][2] * Fz[k];
        uy.a[k] = (B[1][0] * Fx[k] + B[1][1] * Fy[k]) + B[1][2] * Fz[k];
        uz.a[k] = (B[2][0] * Fx[k] + B[2][1] * Fy[k]) + B[2][2] * Fz[k];
    }
    double Lx[kMG] = {0}, Ly[kMG] = {0}, Lz[kMG] = {0};
    for (int q = 0; q < nf; ++q) {
        const FaceGeo& g = fg[q];
        for (int k = 0; k < kMG; ++k) {
            const double dF_p = mg_comp(g.zn, k) - mg_comp(zc, k);
            const double du_n = (ux.a[k] * g.dr.x + uy.a[k] * g.dr.y) + uz.a[k] * g.dr.z;
            const double dF_m = 2.0 * du_n - dF_p;
            const double dF_lim = lim_binary(S.limiter, dF_m, dF_p) * dF_p;
            Lx[k] = Lx[k] + (g.w * dF_lim) * g.drn.x;
            Ly[k] = Ly[k] + (g.w * dF_lim) * g.drn.y;
            Lz[k] = Lz[k] + (g.w * dF_lim) * g.drn.z;
        }
    }
    for (int k = 0; k < kMG; ++k) {
        gx.a[k] = (B[0][0] * Lx[k] + B[0][1] * Ly[k]) + B[0][2] * Lz[k];
        gy.a[k] = (B[1][0] * Lx[k] + B[1][1] * Ly[k]) + B[1][2] * Lz[k];
        gz.a[k] = (B[2][0] * Lx[k] + B[2][1] * Ly[k]) + B[2][2] * Lz[k];
    }
}

// PState::is_bad, mmf (models.cpp:393-403); velocity z is zero.
inline bool m_is_bad(const MState& z) {
    bool empty = true;   // Fractions::empty (fractions.cpp:91-99)
    for (int i = 0; i < kMaxMat; ++i)
        if (z.beta.a[i] > std::numeric_limits<double>::epsilon() / 4) empty = false;
    auto bad = [](double x) { return std::isinf(x) || std::isnan(x); };
    return bad(z.rho) || bad(z.v.x) || bad(z.v.y) || bad(z.v.z) || bad(z.P) || bad(z.e) || bad(z.T) || empty ||
           z.rho < 0;
}

// MmFluid::distributor (mm_fluid.cpp:765-830) with the storage moves of
// an adaptation.  `bad`: a child or parent state the reference rejects (it
// throws "bad cell").
struct MmData : AmrData {
    orc_mm* S;
    bool bad = false;
    explicit MmData(orc_mm* s) : S(s) {}
    void resize(int n) override {
        S->init.resize(n);
        S->half.resize(n);
        S->next.resize(n);
        S->d_dx.resize(n);
        S->d_dy.resize(n);
        S->d_dz.resize(n);
    }
    // split (:768-800): child = parent + grad . dr with momentum- and
    // energy-consistent corrections, beta normalised, P from
    // pressure_re(rho, e, beta, {P0, T0} of the parent), T from
    // temperature_rP(rho, P, beta, {T0}); the densities stay the
    // default-constructed ScalarSet (zeros).
    void split(const orc_mesh& m, int ip, const int* ch) override {
        const MState zp = S->init[ip];
        const MG gx = S->d_dx[ip], gy = S->d_dy[ip], gz = S->d_dz[ip];
        for (int k = 0; k < 4; ++k) {
            const int c = ch[k];
            const Vec3 dr = Vec3{m.cx[c], m.cy[c], 0.0} - Vec3{m.cx[ip], m.cy[ip], 0.0};
            double delta[kMG];
            for (int j = 0; j < kMG; ++j) delta[j] = (gx.a[j] * dr.x + gy.a[j] * dr.y) + gz.a[j] * dr.z;
            MState zc;
            zc.rho = zp.rho + delta[0];
            const double ratio = zp.rho / zc.rho;
            zc.v = {zp.v.x + ratio * delta[1], zp.v.y + ratio * delta[2], zp.v.z + ratio * 0.0};
            for (int i = 0; i < S->mix.n; ++i) zc.beta.a[i] = zp.beta.a[i] + (zp.rho / zc.rho) * delta[4 + i];
            zc.beta.normalize();
            const Vec3 dv{zp.v.x - zc.v.x, zp.v.y - zc.v.y, zp.v.z - zc.v.z};
            zc.e = (zp.e + 0.5 * sqnorm(dv)) + (zp.rho / zc.rho) * delta[3];
            double rh[kMaxMat];
            MixOpt op;
            op.P0 = zp.P;
            op.T0 = zp.T;
            const int idx = zc.beta.index();
            double Tt;
            if (idx >= 0) zc.P = sg_pressure_re(S->mix.m[idx], zc.rho, zc.e);   // MixturePT::pressure_re (:26-36)
            else find_rPT(S->mix, zc.rho, zc.e, zc.beta, op, rh, zc.P, Tt);
            MixOpt ot;
            ot.T0 = zp.T;
            if (idx >= 0) zc.T = sg_temperature_rP(S->mix.m[idx], zc.rho, zc.P);   // temperature_rP (:74-84)
            else {
                double TdR, TdP;
                find_rT_rP(S->mix, zc.rho, zc.P, zc.beta, ot, rh, zc.T, TdR, TdP);
            }
            for (int i = 0; i < kMaxMat; ++i) zc.rhos[i] = 0.0;
            if (m_is_bad(zc)) bad = true;
            S->init[c] = zc;
        }
    }
    // merge (:802-830): volume-weighted conserved sum, mean P and T as the
    // closure's starting point, negative partial densities clipped.
    void merge(const orc_mesh& m, const int* ch, int ip) override {
        MArr sum;
        double mean_p = 0.0, mean_t = 0.0;
        for (int k = 0; k < 4; ++k) {
            const MState& zc = S->init[ch[k]];
            const MArr qc = m_qstate(zc);
            const double V = m.vol[ch[k]];
            for (int j = 0; j < 9; ++j) sum.a[j] = sum.a[j] + qc.a[j] * V;
            mean_p += zc.P * V;
            mean_t += zc.T * V;
        }
        for (int j = 0; j < 9; ++j) sum.a[j] = sum.a[j] / m.vol[ip];
        mean_p /= m.vol[ip];
        mean_t /= m.vol[ip];
        for (int i = 0; i < kMaxMat; ++i) sum.a[5 + i] = std::max(0.0, sum.a[5 + i]);
        const MState z = m_from_q(sum, S->mix, mean_p, mean_t, nullptr);
        if (m_is_bad(z)) bad = true;
        S->init[ip] = z;
    }
    void move(int from, int to) override {
        S->init[to] = S->init[from];
        S->d_dx[to] = S->d_dx[from];
        S->d_dy[to] = S->d_dy[from];
        S->d_dz[to] = S->d_dz[from];
    }
};

// AmrCells::hx / hy (mesh/euler/amr_cells.cpp:91-99): |vs<+1,0> - vs<-1,0>|, |vs<0,+1> - vs<0,-1>|.
inline double cell_hx(const orc_mesh& m, int c) {
    const double* X = &m.vx[9 * (size_t)c];
    const double* Y = &m.vy[9 * (size_t)c];
    Vec3 d{X[5] - X[3], Y[5] - Y[3], 0.0};
    return std::sqrt(sqnorm(d));
}
inline double cell_hy(const orc_mesh& m, int c) {
    const double* X = &m.vx[9 * (size_t)c];
    const double* Y = &m.vy[9 * (size_t)c];
    Vec3 d{X[7] - X[1], Y[7] - Y[1], 0.0};
    return std::sqrt(sqnorm(d));
}

// FaceExtra::Direct(...).m / .p (face_extra.h:122-184): every component
// moved by du = (g_x dr.x + g_y dr.y) + g_z dr.z of its own gradient.
MState m_extrap(const MState& z, const MG& gx, const MG& gy, const MG& gz, const Vec3& dr) {
    double d[kMG];
    for (int k = 0; k < kMG; ++k) d[k] = (gx.a[k] * dr.x + gy.a[k] * dr.y) + gz.a[k] * dr.z;
    MState r = z;
    r.rho = z.rho + d[0];
    r.v = {z.v.x + d[1], z.v.y + d[2], z.v.z};
    r.e = z.e + d[3];
    for (int i = 0; i < kMaxMat; ++i) r.beta.a[i] = z.beta.a[i] + d[4 + i];
    r.P = z.P + d[8];
    r.T = z.T + d[9];
    for (int i = 0; i < kMaxMat; ++i) r.rhos[i] = z.rhos[i] + d[10 + i];
    return r;
}

// PState::interpolation_update (models.cpp:381-391): normalise the mass
// fractions, then e, T and the true densities from (rho, P) with the
// extrapolated T and densities as the closure's starting point.
void m_interpolation_update(MState& z, const Mixture& mx) {
    z.beta.normalize();
    MixOpt o;
    o.T0 = z.T;
    o.rhos = z.rhos;
    double rh[kMaxMat], e, T;
    get_reT(mx, z.rho, z.P, z.beta, o, rh, e, T);
    z.e = e;
    z.T = T;
    for (int i = 0; i < kMaxMat; ++i) z.rhos[i] = rh[i];
}

// MixturePT::stiffened_gas (mixture_pt.cpp:152-169): the stiffened gas
// tangent to the mixture EOS at (rho, P) -- a pure state's own material,
// else gamma = 1 + 1 / (rho e_P), P0 = -(rho e_R + P e_P) / (gamma e_P)
// with e_R, e_P from find_reT_rP with derivatives (e0 is not needed).
inline void mix_stiffened_gas(const Mixture& mx, double rho, double P, const Frac& b, const MixOpt& o0, double& g,
                              double& p0) {
    const int idx = b.index();
    if (idx >= 0) {   // StiffenedGas::stiffened_gas (stiffened_gas.cpp:169-171)
        g = mx.m[idx].g;
        p0 = mx.m[idx].P0;
        return;
    }
    MixOpt o = o0;
    o.deriv = true;
    double rh[kMaxMat], e, edR, edP, T;
    find_reT_rP(mx, rho, P, b, o, rh, e, edR, edP, T);
    g = 1.0 + 1.0 / (rho * edP);
    p0 = -(rho * edR + P * edP) / (g * edP);
}

// Godunov::calc_flux, mmf (math/cfd/flux/godunov.cpp:39-77): the exact
// Riemann solver between the two sides' tangent stiffened gases; the face
// state takes the upwind side's (U >= 0: left) tangential velocity and
// fractions, its energy and temperature from get_reT at the face's
// (rho, P) with the side's T as the guess; the physical flux of it.
inline MArr m_godunov(const MState& zL, const MState& zR, const Mixture& mx) {
    if (m_is_bad(zL) || m_is_bad(zR)) g_numeric_error = 1;   // :40-47 throw
    MixOpt oL, oR;
    oL.T0 = zL.T;
    oL.rhos = zL.rhos;
    oR.T0 = zR.T;
    oR.rhos = zR.rhos;
    double gL, p0L, gR, p0R;
    mix_stiffened_gas(mx, zL.rho, zL.P, zL.beta, oL, gL, p0L);
    mix_stiffened_gas(mx, zR.rho, zR.P, zR.beta, oR, gR, p0R);
    const RSol sol = riemann_solve(zL.rho, zL.v.x, zL.P, gL, p0L, zR.rho, zR.v.x, zR.P, gR, p0R);
    const MState& zs = sol.U >= 0 ? zL : zR;
    MixOpt os;
    os.T0 = zs.T;
    MState z;
    z.rho = sol.rho;
    z.v = {sol.Uf, zs.v.y, zs.v.z};
    z.P = sol.Pf;
    double rh[kMaxMat];
    get_reT(mx, sol.rho, sol.Pf, zs.beta, os, rh, z.e, z.T);
    z.beta = zs.beta;
    for (int i = 0; i < kMaxMat; ++i) z.rhos[i] = NAN;
    MArr F = m_flux(z);
    for (int k = 0; k < 5; ++k)
        if (!std::isfinite(F.a[k])) g_numeric_error = 1;   // Flux::is_bad, :65-75 exits
    return F;
}

// MmFluid::calc_flux (mm_fluid.cpp:347-371) with alpha_sigma for
// CrpMode::PLIC (:228-256): face-frame states z_m / z_p of slot f of cell
// c; the face-frame flux.  inx / iny: the interface normals of
// interface_recovery, [mat][n].
MArr mm_calc_flux(const orc_mm& S, int c, size_t f, const MState& z_m, const MState& z_p, double h_L, double h_R,
                  double dt, const double* inx, const double* iny, int& crp_err) {
    const orc_mesh& m = *S.m;
    if (!S.crp || (z_m.beta.index() >= 0 && z_p.beta.index() >= 0)) {
        if (S.flux == kFluxGodunov) return m_godunov(z_m, z_p, S.mix);
        if (S.flux == kFluxRusanov) return m_rusanov(z_m, z_p, S.mix);
        if (S.flux == kFluxHll) return m_hll(z_m, z_p, S.mix);
        return m_hllc(z_m, z_p, S.mix);
    }
    const int iA = master_component(z_m, z_p);
    const double a_L = m_alpha(z_m, iA), a_R = m_alpha(z_p, iA);
    if (S.crp != kCrpPlic) {   // the heuristic templates (alpha_sigma, :235-243)
        const double a_sig = S.crp == kCrpV3 ? face_fraction_v3(a_L, a_R)
                           : S.crp == kCrpV5 ? face_fraction_v5(a_L, a_R) : face_fraction_s(a_L, a_R);
        return calc_crp_flux(z_m, z_p, h_L, h_R, iA, a_sig, dt, S.mix, crp_err);
    }
    const double vi = 0.5 * (z_m.v.x + z_p.v.x);
    const bool upwind = vi > 0.0;
    const double CFL = dt * std::abs(vi) / (upwind ? h_L : h_R);
    const int cn = slot_is_boundary(m.bnd[f]) ? c : m.adj[f];   // face.neib() is the cell at a boundary
    const size_t ki = (size_t)iA * m.n;
    const Vec3 normal{m.fnx[f], m.fny[f], 0.0};
    const Vec3 nc = upwind ? Vec3{inx[ki + c], iny[ki + c], 0.0} : Vec3{-inx[ki + cn], -iny[ki + cn], -0.0};
    const double cs = dot(normal, nc);
    const double a_sig = f_between(average_flux(upwind ? a_L : a_R, cs, CFL), a_L, a_R);
    return calc_crp_flux(z_m, z_p, h_L, h_R, iA, a_sig, dt, S.mix, crp_err);
}

// EuCell::faces(dir) (amr_faces.cpp:186-197): 1 = X keeps faces with
// |n.x| >= 0.7, 2 = Y those with |n.y| >= 0.7, 0 = all.
inline bool mm_skip(const orc_mesh& m, size_t f, int dir) {
    if (dir == 1) return std::abs(m.fnx[f]) < 0.7;
    if (dir == 2) return std::abs(m.fny[f]) < 0.7;
    return false;
}

inline void m_to_global(MArr& loc, const Vec3& normal) {   // Flux::to_global
    const Vec3 mg = rot_global(Vec3{loc.a[1], loc.a[2], loc.a[3]}, normal);
    loc.a[1] = mg.x;
    loc.a[2] = mg.y;
    loc.a[3] = mg.z;
}

}  // namespace

extern "C" {

orc_mm* orc_mm_create(orc_mesh* m, int n_mat, const double* mats, double cfl) {
    orc_mm* s = new orc_mm();
    s->m = m;
    s->mix.n = n_mat;
    for (int i = 0; i < n_mat; ++i)
        s->mix.m[i] = SgMat{mats[6 * i], mats[6 * i + 1], mats[6 * i + 2], mats[6 * i + 3], mats[6 * i + 4],
                            mats[6 * i + 5]};
    s->cfl = std::max(0.0, std::min(cfl, 1.0));   // set_CFL (mm_fluid.cpp:41-43)
    s->dt = 1.0e+300;
    s->max_dt = 1.0e+300;                          // mm_fluid.cpp:16-26
    s->init.resize(m->n);
    s->half.resize(m->n);
    s->next.resize(m->n);
    s->d_dx.resize(m->n);
    s->d_dy.resize(m->n);
    s->d_dz.resize(m->n);
    return s;
}

void orc_mm_set_limiter(orc_mm* s, int limiter) { s->limiter = limiter; }   // set_limiter (mm_fluid.cpp:45-47)
void orc_mm_set_accuracy(orc_mm* s, int acc) { s->acc = std::min(std::max(1, acc), 2); }   // mm_fluid.cpp:46-48

// MmFluid::compute_grad (mm_fluid.cpp:439-466) of the init layer.
void orc_mm_compute_grad(orc_mm* s) {
    const int n = s->m->n;
    #pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int c = 0; c < n; ++c) mm_grad_cell(*s, c, s->d_dx[c], s->d_dy[c], s->d_dz[c]);
}

void orc_mm_set_threads(orc_mm* s, int nthreads) { s->nthreads = std::max(1, nthreads); }

// The stored gradients: rho, u, v, e, beta[n_mat], P, T, rhos[n_mat]
// rows of n values each, x then y.
void orc_mm_grad(const orc_mm* s, double* gx, double* gy) {
    const int n = s->m->n, nm = s->mix.n, K = 6 + 2 * nm;
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < K; ++k) {
            const int j = k < 4 + nm ? k : (k < 6 + nm ? 8 + (k - 4 - nm) : 10 + (k - 6 - nm));
            gx[(size_t)k * n + c] = s->d_dx[c].a[j];
            gy[(size_t)k * n + c] = s->d_dy[c].a[j];
        }
}

// MmFluid::set_flags (mm_fluid.cpp:836-885): compute_grad, then +1 for a
// mixed cell or a pure cell with a face neighbour of another material (or
// mixed; boundary faces skipped), else -1.
void orc_mm_set_flags(orc_mm* s) {
    orc_mm_compute_grad(s);
    orc_mesh& m = *s->m;
    for (int c = 0; c < m.n; ++c) {
        const MState& z = s->init[c];
        if (!z.beta.is_pure()) {
            m.flag[c] = 1;
            continue;
        }
        const int idx = z.beta.index();
        bool need_split = false;
        for (int s8 = 0; s8 < 8 && !need_split; ++s8) {
            const size_t f = 8 * (size_t)c + s8;
            const int b = m.bnd[f];
            if (b == ORC_UNDEFINED || slot_is_boundary(b)) continue;
            if (s->init[m.adj[f]].beta.index() != idx) need_split = true;
        }
        m.flag[c] = need_split ? 1 : -1;
    }
}

// EuMesh::refine (eu_mesh.cpp:290-359) with MmFluid::distributor: 1 when
// the mesh changed, 0 when not, -1 when a split or merge gave a state the
// reference rejects (it throws).
int orc_mm_refine(orc_mm* s, int max_level) {
    balance_flags(*s->m, max_level);
    MmData D(s);
    const bool changed = apply_flags(*s->m, &D);
    if (D.bad) return -1;
    return changed ? 1 : 0;
}

void orc_mm_free(orc_mm* s) { delete s; }

void orc_mm_set_max_dt(orc_mm* s, double v) { s->max_dt = v; }

// Cells from {rho, u, v, P, beta[mat][n]}: the problem files' PState(rho, v,
// P, Fractions{...}, mixture) (models.cpp:176-185; the Fractions constructor
// normalises, fractions.cpp:6-27).
void orc_mm_set_state(orc_mm* s, const double* rho, const double* u, const double* v, const double* P,
                      const double* beta) {
    const int n = s->m->n;
    for (int c = 0; c < n; ++c) {
        Frac b;
        for (int i = 0; i < s->mix.n; ++i) b.a[i] = beta[(size_t)i * n + c];
        b.normalize();
        s->init[c] = m_from_rvp(s->mix, rho[c], u[c], v[c], P[c], b);
    }
}

void orc_mm_state(const orc_mm* s, double* rho, double* u, double* v, double* P, double* e, double* T,
                  double* beta, double* rhos) {
    const int n = s->m->n;
    for (int c = 0; c < n; ++c) {
        const MState& z = s->init[c];
        rho[c] = z.rho; u[c] = z.v.x; v[c] = z.v.y; P[c] = z.P; e[c] = z.e; T[c] = z.T;
        for (int i = 0; i < s->mix.n; ++i) {
            beta[(size_t)i * n + c] = z.beta.a[i];
            rhos[(size_t)i * n + c] = z.rhos[i];
        }
    }
}

// MmFluid::compute_dt (mm_fluid.cpp:166-194).
double orc_mm_compute_dt(orc_mm* s) {
    const orc_mesh& m = *s->m;
    // per-cell candidates in parallel, then EuMesh::min over them in cell order
    std::vector<double> cand(m.n);
    int nerr = 0;
    const int err0 = g_numeric_error;
    #pragma omp parallel for num_threads(s->nthreads) schedule(static) reduction(|:nerr)
    for (int c = 0; c < m.n; ++c) {
        g_numeric_error = 0;
        double dt = std::numeric_limits<double>::max();
        const MState& z = s->init[c];
        MixOpt o;
        o.T0 = z.T;
        o.rhos = z.rhos;
        const double cs = mix_sound_speed_rP(s->mix, z.rho, z.P, z.beta, o);
        for (int k = 0; k < 8; ++k) {
            const size_t f = 8 * (size_t)c + k;
            if (m.bnd[f] == ORC_UNDEFINED) continue;
            const Vec3 n{m.fnx[f], m.fny[f], 0.0};
            const double vn = dot(z.v, n);
            const double lambda = std::abs(vn) + cs;
            dt = smin(dt, m.vol[c] / (m.area[f] * lambda));
        }
        cand[c] = dt;
        nerr |= g_numeric_error;
    }
    g_numeric_error = err0 | nerr;
    double dtmin = std::numeric_limits<double>::max();
    for (int c = 0; c < m.n; ++c) dtmin = smin(dtmin, cand[c]);   // EuMesh::min
    s->dt = smin(s->cfl * dtmin, s->max_dt);
    return s->dt;
}

void orc_mm_set_crp(orc_mm* s, int mode) { s->crp = mode; }   // CrpMode value (0 NONE .. 4 PLIC)
void orc_mm_set_flux(orc_mm* s, int flux) { s->flux = flux; }  // Fluxes value of the non-composite faces
void orc_mm_set_splitting(orc_mm* s, int split) { s->split = split; }

// MmFluid::update, first order (mm_fluid.cpp:110-145): compute_dt, for
// CrpMode::PLIC interface_recovery, fluxes (:373-445) with calc_flux
// (:339-371: HLLC between two pure states or without CRP, else the
// composite Riemann flux of the master material with the PLIC face
// fraction, alpha_sigma :228-256), swap; returns 0, or 1 on a NaN flux, a
// negative density or a CRP case the reference rejects (it throws).
// MmFluid::integrate with acc 1 (mm_fluid.cpp:139-150): interface_recovery
// (PLIC), fluxes (:373-445) over the faces of direction dir, swap.
static int mm_integrate1(orc_mm* s, double dt, int dir) {
    const orc_mesh& m = *s->m;
    const int nm = s->mix.n;
    std::vector<double> inx, iny, ip;
    if (s->crp == kCrpPlic) {
        inx.resize((size_t)nm * m.n);
        iny.resize((size_t)nm * m.n);
        ip.resize((size_t)nm * m.n);
        orc_mm_interface(s, inx.data(), iny.data(), ip.data());
    }
    int neg = 0, crp_err = 0, nerr = 0;
    const int err0 = g_numeric_error;   // the loop resets this thread's copy per cell
    #pragma omp parallel for num_threads(s->nthreads) schedule(static) reduction(|:neg, crp_err, nerr)
    for (int c = 0; c < m.n; ++c) {
        g_numeric_error = 0;   // thread-local; gathered below
        const double V_c = m.vol[c];
        const MState z_c = s->init[c];
        MArr flux;
        for (int k = 0; k < 8; ++k) {
            const size_t f = 8 * (size_t)c + k;
            const int b = m.bnd[f];
            if (b == ORC_UNDEFINED || mm_skip(m, f, dir)) continue;
            const Vec3 normal{m.fnx[f], m.fny[f], 0.0};
            MState z_n;
            if (!slot_is_boundary(b)) z_n = s->init[m.adj[f]];
            else z_n = m_boundary_value(z_c, normal, b);
            const MState z_m = m_in_local(z_c, normal);
            const MState z_p = m_in_local(z_n, normal);
            const double S = m.area[f];
            const double h_L = V_c / S;
            const double h_R = (slot_is_boundary(b) ? V_c : m.vol[m.adj[f]]) / S;
            MArr loc = mm_calc_flux(*s, c, f, z_m, z_p, h_L, h_R, dt, inx.data(), iny.data(), crp_err);
            m_to_global(loc, normal);
            for (int q = 0; q < 9; ++q) flux.a[q] += loc.a[q] * S;
        }
        MArr q_c = m_qstate(z_c);
        const double coef = dt / V_c;
        for (int q = 0; q < 9; ++q) q_c.a[q] -= coef * flux.a[q];
        if (q_c.a[0] < 0.0) neg = 1;
        s->next[c] = m_from_q(q_c, s->mix, z_c.P, z_c.T, z_c.rhos);
        nerr |= g_numeric_error;
    }
    g_numeric_error = err0 | nerr;
    std::swap(s->init, s->next);   // MmFluid::swap (mm_fluid.cpp:682-684)
    return (neg || crp_err) ? 1 : 0;
}

// MmFluid::integrate with acc 2 (mm_fluid.cpp:153-170): compute_grad of
// init, interface_recovery (PLIC), fluxes_stage1 (:521-578) into half and
// fluxes_stage2 (:580-680) into next over the faces of direction dir, swap.
static int mm_integrate2(orc_mm* s, double dt, int dir) {
    const orc_mesh& m = *s->m;
    const int nm = s->mix.n;
    orc_mm_compute_grad(s);
    std::vector<double> inx, iny, ip;
    if (s->crp == kCrpPlic) {
        inx.resize((size_t)nm * m.n);
        iny.resize((size_t)nm * m.n);
        ip.resize((size_t)nm * m.n);
        orc_mm_interface(s, inx.data(), iny.data(), ip.data());
    }
    // stage 1: mixed cells keep their state; a pure cell advances half a
    // step with the physical flux of its own extrapolated face states
    int nerr = 0;
    const int err0 = g_numeric_error;   // the loops reset this thread's copy per cell
    #pragma omp parallel for num_threads(s->nthreads) schedule(static) reduction(|:nerr)
    for (int c = 0; c < m.n; ++c) {
        g_numeric_error = 0;
        const MState z_c = s->init[c];
        if (z_c.beta.index() < 0) {
            s->half[c] = z_c;
            continue;
        }
        const Vec3 cell_c{m.cx[c], m.cy[c], 0.0};
        MArr flux;
        for (int k = 0; k < 8; ++k) {
            const size_t f = 8 * (size_t)c + k;
            if (m.bnd[f] == ORC_UNDEFINED || mm_skip(m, f, dir)) continue;
            const Vec3 normal{m.fnx[f], m.fny[f], 0.0};
            const Vec3 face_c{m.fcx[f], m.fcy[f], 0.0};
            MState z_m = m_extrap(z_c, s->d_dx[c], s->d_dy[c], s->d_dz[c], face_c - cell_c);
            m_interpolation_update(z_m, s->mix);
            MArr loc = m_flux(m_in_local(z_m, normal));
            m_to_global(loc, normal);
            for (int q = 0; q < 9; ++q) flux.a[q] += loc.a[q] * m.area[f];
        }
        MArr q_c = m_qstate(z_c);
        const double coef = 0.5 * dt / m.vol[c];
        for (int q = 0; q < 9; ++q) q_c.a[q] -= coef * flux.a[q];
        s->half[c] = m_from_q(q_c, s->mix, z_c.P, z_c.T, z_c.rhos);
        nerr |= g_numeric_error;
    }
    // stage 2: face states extrapolated from the half states with the
    // gradients of init; a periodic face's neighbour centre is the
    // neighbour's own (face.is_boundary() is false there, :599-604)
    int neg = 0, crp_err = 0;
    #pragma omp parallel for num_threads(s->nthreads) schedule(static) reduction(|:neg, crp_err, nerr)
    for (int c = 0; c < m.n; ++c) {
        g_numeric_error = 0;
        const Vec3 cell_c{m.cx[c], m.cy[c], 0.0};
        const double V_c = m.vol[c];
        const MState z_c = s->init[c];
        const MState z_ch = s->half[c];
        MArr flux;
        for (int k = 0; k < 8; ++k) {
            const size_t f = 8 * (size_t)c + k;
            const int b = m.bnd[f];
            if (b == ORC_UNDEFINED || mm_skip(m, f, dir)) continue;
            const Vec3 normal{m.fnx[f], m.fny[f], 0.0};
            const Vec3 face_c{m.fcx[f], m.fcy[f], 0.0};
            const bool bound = slot_is_boundary(b);
            const int nb = m.adj[f];   // the cell itself on boundary slots
            MState z_m = m_extrap(z_ch, s->d_dx[c], s->d_dy[c], s->d_dz[c], face_c - cell_c);
            m_interpolation_update(z_m, s->mix);
            MState z_p;
            double V_n = V_c;
            if (!bound) {
                const Vec3 neib_c{m.cx[nb], m.cy[nb], 0.0};
                V_n = m.vol[nb];
                z_p = m_extrap(s->half[nb], s->d_dx[nb], s->d_dy[nb], s->d_dz[nb], face_c - neib_c);
                m_interpolation_update(z_p, s->mix);
            } else {
                z_p = m_boundary_value(z_m, normal, b);
            }
            const double S = m.area[f];
            MArr loc = mm_calc_flux(*s, c, f, m_in_local(z_m, normal), m_in_local(z_p, normal), V_c / S, V_n / S, dt,
                                    inx.data(), iny.data(), crp_err);
            m_to_global(loc, normal);
            for (int q = 0; q < 9; ++q) flux.a[q] += loc.a[q] * S;
        }
        MArr q_c = m_qstate(z_c);
        const double coef = dt / V_c;
        for (int q = 0; q < 9; ++q) q_c.a[q] -= coef * flux.a[q];
        if (q_c.a[0] < 0.0) neg = 1;   // :671-673 throws
        s->next[c] = m_from_q(q_c, s->mix, z_c.P, z_c.T, z_c.rhos);
        nerr |= g_numeric_error;
    }
    g_numeric_error = err0 | nerr;
    std::swap(s->init, s->next);
    return (neg || crp_err) ? 1 : 0;
}

// MmFluid::update (mm_fluid.cpp:112-137): compute_dt, then integrate over
// all faces, or X then Y (SIMPLE), or X for dt / 2, Y for dt, X for dt / 2
// (STRANG).
int orc_mm_update(orc_mm* s) {
    g_numeric_error = 0;
    orc_mm_compute_dt(s);
    const double dt = s->dt;
    auto integrate = [s](double h, int dir) { return s->acc == 2 ? mm_integrate2(s, h, dir) : mm_integrate1(s, h, dir); };
    int r = 0;
    if (s->split == 1) {
        r |= integrate(dt, 1);
        r |= integrate(dt, 2);
    } else if (s->split == 2) {
        r |= integrate(0.5 * dt, 1);
        r |= integrate(dt, 2);
        r |= integrate(0.5 * dt, 1);
    } else {
        r = integrate(dt, 0);
    }
    return (g_numeric_error || r) ? 1 : 0;
}

double orc_mm_dt(const orc_mm* s) { return s->dt; }

// MmFluid::interface_recovery (mm_fluid.cpp:484-519): normals nx/ny and
// section positions p per material [mat][n] (zero for pure cells and
// absent materials).
void orc_mm_interface(orc_mm* s, double* nx, double* ny, double* p) {
    const orc_mesh& m = *s->m;
    const int n = m.n, nm = s->mix.n;
    #pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int c = 0; c < n; ++c) {
        Vec3 ns[kMaxMat];
        double ps[kMaxMat];
        for (int i = 0; i < nm; ++i) {
            ns[i] = {0.0, 0.0, 0.0};
            ps[i] = 0.0;
        }
        const MState& z = s->init[c];
        if (!z.beta.is_pure()) {
            const Frac a_c = m_volume_fractions(z);
            for (int k = 0; k < 8; ++k) {
                const size_t f = 8 * (size_t)c + k;
                const int b = m.bnd[f];
                if (b == ORC_UNDEFINED) continue;
                const Vec3 S = scale(m.area[f], Vec3{m.fnx[f], m.fny[f], 0.0});
                const MState& zn = slot_is_boundary(b) ? z : s->init[m.adj[f]];   // neib() = the cell at boundaries
                for (int i = 0; i < nm; ++i)
                    if (a_c.has(i)) {
                        const double a_n = m_alpha(zn, i);
                        const Vec3 t = scale(face_fraction(a_c.a[i], a_n), S);
                        ns[i] = {ns[i].x - t.x, ns[i].y - t.y, ns[i].z - t.z};
                    }
            }
            for (int i = 0; i < nm; ++i)
                if (a_c.has(i)) {
                    const double zz = sqnorm(ns[i]);   // Vector3d::normalize
                    if (zz > 0.0) {
                        const double r = std::sqrt(zz);
                        ns[i] = {ns[i].x / r, ns[i].y / r, ns[i].z / r};
                    }
                    ps[i] = quad_find_section(a_c.a[i], ns[i].x, ns[i].y, cell_hx(m, c), cell_hy(m, c));
                }
        }
        for (int i = 0; i < nm; ++i) {
            nx[(size_t)i * n + c] = ns[i].x;
            ny[(size_t)i * n + c] = ns[i].y;
            p[(size_t)i * n + c] = ps[i];
        }
    }
}

// Scalar entry points for the known-answer tests.
void orc_mm_closure(int n_mat, const double* mats, double rho, double e, const double* beta, double P0, double T0,
                    double* out /* P, T, rhos[4] */) {
    Mixture mx;
    mx.n = n_mat;
    for (int i = 0; i < n_mat; ++i)
        mx.m[i] = SgMat{mats[6 * i], mats[6 * i + 1], mats[6 * i + 2], mats[6 * i + 3], mats[6 * i + 4], mats[6 * i + 5]};
    Frac b;
    for (int i = 0; i < n_mat; ++i) b.a[i] = beta[i];
    MixOpt o;
    o.P0 = P0;
    o.T0 = T0;
    get_rPT(mx, rho, e, b, o, out + 2, out[0], out[1]);
}

void orc_mm_closure_rP(int n_mat, const double* mats, double rho, double P, const double* beta,
                       double* out /* e, T, c, rhos[4] */) {
    Mixture mx;
    mx.n = n_mat;
    for (int i = 0; i < n_mat; ++i)
        mx.m[i] = SgMat{mats[6 * i], mats[6 * i + 1], mats[6 * i + 2], mats[6 * i + 3], mats[6 * i + 4], mats[6 * i + 5]};
    Frac b;
    for (int i = 0; i < n_mat; ++i) b.a[i] = beta[i];
    MixOpt o;
    get_reT(mx, rho, P, b, o, out + 3, out[0], out[1]);
    MixOpt oc;
    oc.T0 = out[1];
    oc.rhos = out + 3;
    out[2] = mix_sound_speed_rP(mx, rho, P, b, oc);
}

// mmf HLLC between two cells of an orc_mm state (face frame = normal (nx, ny)).
void orc_mm_face_flux(const orc_mm* s, int cL, int cR, double nx, double ny, double* F) {
    const Vec3 n{nx, ny, 0.0};
    MArr f = m_hllc(m_in_local(s->init[cL], n), m_in_local(s->init[cR], n), s->mix);
    for (int k = 0; k < 9; ++k) F[k] = f.a[k];
}

// The plain mmf flux `kind` (0 Godunov, 1 Rusanov, 2 HLL, 3 HLLC) between
// the states of cells cL and cR across a face with normal (nx, ny), in the
// face frame; returns 1 where the reference throws.
int orc_mm_face_flux_kind(const orc_mm* s, int kind, int cL, int cR, double nx, double ny, double* F) {
    const Vec3 n{nx, ny, 0.0};
    const MState zL = m_in_local(s->init[cL], n), zR = m_in_local(s->init[cR], n);
    g_numeric_error = 0;
    MArr f = kind == kFluxGodunov ? m_godunov(zL, zR, s->mix)
           : kind == kFluxRusanov ? m_rusanov(zL, zR, s->mix)
           : kind == kFluxHll ? m_hll(zL, zR, s->mix) : m_hllc(zL, zR, s->mix);
    for (int k = 0; k < 9; ++k) F[k] = f.a[k];
    return g_numeric_error;
}

// MixturePT::stiffened_gas of cell c's state: out = {gamma, P0}.
void orc_mm_stiffened_gas(const orc_mm* s, int c, double* out) {
    const MState& z = s->init[c];
    MixOpt o;
    o.T0 = z.T;
    o.rhos = z.rhos;
    mix_stiffened_gas(s->mix, z.rho, z.P, z.beta, o, out[0], out[1]);
}

// sound_speed_rP of cell c's state.
double orc_mm_sound_speed(const orc_mm* s, int c) {
    const MState& z = s->init[c];
    MixOpt o;
    o.T0 = z.T;
    o.rhos = z.rhos;
    return mix_sound_speed_rP(s->mix, z.rho, z.P, z.beta, o);
}

// calc_crp_flux (mm_fluid.cpp:259-337) between the states of cells cL and
// cR across a face with normal +e_x (the face frame is the global one):
// cell widths hL, hR, master material iA, face fraction a_sig, time step
// dt; F: the 9 flux components; returns the "strange case" error flag.
int orc_mm_crp_flux(const orc_mm* s, int cL, int cR, double hL, double hR, int iA, double a_sig, double dt,
                    double* F) {
    int err = 0;
    const MArr f = calc_crp_flux(s->init[cL], s->init[cR], hL, hR, iA, a_sig, dt, s->mix, err);
    for (int k = 0; k < 9; ++k) F[k] = f.a[k];
    return err;
}
// wave_config (hllc.cpp:256-335) of the states of cells cL and cR in the
// same frame: {S_L, S_C, S_R}.
void orc_mm_wave_speeds(const orc_mm* s, int cL, int cR, double* S3) {
    const MState& zL = s->init[cL];
    const MState& zR = s->init[cR];
    const Wave w = wave_config(s->mix, m_qstate(zL), m_flux(zL), m_qstate(zR), m_flux(zR));
    S3[0] = w.SL;
    S3[1] = w.SC;
    S3[2] = w.SR;
}
// CrpFlux::classic (inverse = 0) or CrpFlux::inverse (inverse = 1) of the
// states of cells c0, c1, c2 in that argument order (crp_flux.cpp:125-512),
// across a +x face, with the distance delta and time step dt.
int orc_mm_crp_sub(const orc_mm* s, int inverse, int c0, int c1, int c2, double delta, double dt, double* F) {
    int err = 0;
    const MArr f = inverse ? crp_inverse(s->init[c0], s->init[c1], s->init[c2], s->mix, delta, dt, err)
                           : crp_classic(s->init[c0], s->init[c1], s->init[c2], s->mix, delta, dt, err);
    for (int k = 0; k < 9; ++k) F[k] = f.a[k];
    return err;
}
// PState::split (models.cpp:261-335) of cell c's state for material iA:
// out[0..1] = {rho, e} of the pure part, out[2..3] = {rho, e} of the
// mixture part, out[4..7] its mass fractions, out[8..9] the P and T both
// parts carry.
void orc_mm_split(const orc_mm* s, int c, int iA, double* out) {
    MState zA, zB;
    m_split(s->init[c], s->mix, iA, zA, zB);
    out[0] = zA.rho;
    out[1] = zA.e;
    out[2] = zB.rho;
    out[3] = zB.e;
    for (int i = 0; i < kMaxMat; ++i) out[4 + i] = zB.beta.a[i];
    out[8] = zB.P;
    out[9] = zB.T;
}
void orc_mm_plant(int k) { g_orc_plant = k; }

double orc_quad_volume_fraction(double p, double nx, double ny, double A, double B) {
    return quad_volume_fraction(p, nx, ny, A, B);
}
double orc_quad_find_section(double alpha, double nx, double ny, double A, double B) {
    return quad_find_section(alpha, nx, ny, A, B);
}
double orc_face_fraction(int kind, double a1, double a2) {
    switch (kind) {
        case 3: return face_fraction_v3(a1, a2);
        case 5: return face_fraction_v5(a1, a2);
        case 1: return face_fraction_s(a1, a2);
        default: return face_fraction(a1, a2);
    }
}
double orc_average_flux(double alpha, double cs, double CFL) { return average_flux(alpha, cs, CFL); }

}  // extern "C"
