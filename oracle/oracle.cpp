// ORACLE — test infrastructure only (see oracle.h).  CPU restatement of
// zephyr's SmFluid step on AMR-2D meshes, written from the reference's
// formulas; compiled with -O3 -ffp-contract=off (no FMA contraction, the
// reference's x86-64 build has no -march, CMakeLists.txt:16-19).
//
// ORDER notes.  Eigen expressions in the reference are restated with one
// fixed order:
//   * element-wise array arithmetic: left to right per element, exactly as
//     the C++ expression reads (Eigen array ops are element-wise);
//   * dot / squaredNorm of a 3-vector: (a0*b0 + a1*b1) + a2*b2;
//   * Matrix3d * Vector3d, row i: (R(i,0)*v0 + R(i,1)*v1) + R(i,2)*v2;
//   * Matrix3d::inverse: Eigen's cofactor/adjugate formula, det as
//     (c00*m00 + c10*m10) + c20*m20, B = adj * (1/det).
// On AMR-2D meshes with axis-aligned normals and z = 0 every such sum has at
// most one non-zero term, so the order changes no value (only, possibly, the
// sign of an exact zero).
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>
#include <omp.h>

namespace {

// ---------------------------------------------------------------- vectors
struct Vec3 {
    double x, y, z;
};

inline Vec3 operator-(const Vec3& a, const Vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator+(const Vec3& a, const Vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 scale(double s, const Vec3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(const Vec3& a, const Vec3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline double sqnorm(const Vec3& a) { return (a.x * a.x + a.y * a.y) + a.z * a.z; }
inline double norm(const Vec3& a) { return std::sqrt(sqnorm(a)); }

// std::min / std::max semantics (return the first argument on ties).
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------- states
// smf::PState {rho, v(3), p, e} (reference math/cfd/models.h:24-28).
struct PState {
    double rho;
    Vec3 v;
    double p;
    double e;
    double* arr() { return &rho; }
    const double* arr() const { return &rho; }
};
static_assert(sizeof(PState) == 6 * sizeof(double), "PState layout");

// smf::QState / smf::Flux {rho, m(3), E} (models.h:80-83, 120-123).
struct QState {
    double rho;
    Vec3 m;
    double E;
    double* arr() { return &rho; }
    const double* arr() const { return &rho; }
};
typedef QState Flux;

// Stiffened-gas EOS (phys/matter/eos/stiffened_gas.cpp:103-175).  IdealGas
// (ideal_gas.cpp:32-104) is the case P0 = e0 = 0: each ideal-gas formula is
// the stiffened one with exact "+0"/"-0" terms, so the values are identical.
struct Eos {
    double gamma, P0, e0;
    // stiffened_gas.cpp:103-104
    double pressure_re(double rho, double e) const { return (gamma - 1.0) * rho * (e - e0) - gamma * P0; }
    // stiffened_gas.cpp:121-122
    double energy_rP(double rho, double P) const { return e0 + (P + gamma * P0) / ((gamma - 1.0) * rho); }
    // stiffened_gas.cpp:147-148
    double sound_speed_rP(double rho, double P) const { return std::sqrt(gamma * (P + P0) / rho); }
    // stiffened_gas.cpp:173-174
    double min_pressure() const { return -P0; }
};

// PState(q, eos) (math/cfd/models.cpp:27-32).
inline PState to_pstate(const QState& q, const Eos& eos) {
    PState z;
    z.rho = q.rho;
    z.v = {q.m.x / z.rho, q.m.y / z.rho, q.m.z / z.rho};
    z.e = q.E / z.rho - 0.5 * sqnorm(z.v);
    z.p = eos.pressure_re(z.rho, z.e);
    return z;
}

// QState(z) (models.cpp:85-89).
inline QState to_qstate(const PState& z) {
    QState q;
    q.rho = z.rho;
    q.m = {z.rho * z.v.x, z.rho * z.v.y, z.rho * z.v.z};
    q.E = z.rho * (z.e + 0.5 * sqnorm(z.v));
    return q;
}

// Flux(z), differential flux along x (models.cpp:131-137).
inline Flux phys_flux(const PState& z) {
    Flux f;
    f.rho = z.rho * z.v.x;
    f.m.x = z.rho * z.v.x * z.v.x + z.p;
    f.m.y = z.rho * z.v.x * z.v.y;
    f.m.z = z.rho * z.v.x * z.v.z;
    f.E = (z.rho * (z.e + 0.5 * sqnorm(z.v)) + z.p) * z.v.x;
    return f;
}

// PState::E() (models.h:68).
inline double total_E(const PState& z) { return z.e + 0.5 * sqnorm(z.v); }

// PState::is_bad (models.cpp:58-63).
inline bool is_bad(const PState& z, const Eos& eos) {
    if (!std::isfinite(z.rho) || !std::isfinite(z.p) || !std::isfinite(z.e)) {
        return true;
    }
    return z.rho <= 0.0 || z.p < eos.min_pressure();
}

// ---------------------------------------------------------------- rotation
// Rotate::to_local / to_global, 2D branch no.z == 0 (math/cfd/rotate.cpp:7-53).
inline Vec3 normalized(const Vec3& n) {
    double z = sqnorm(n);
    if (z > 0.0) {
        double s = std::sqrt(z);
        return {n.x / s, n.y / s, n.z / s};
    }
    return n;
}

inline Vec3 rot_local(const Vec3& v, const Vec3& no) {
    Vec3 n = normalized(no);
    double c = n.x, s = n.y;
    // R = [[c, s, 0], [-s, c, 0], [0, 0, 1]]
    return {(c * v.x + s * v.y) + 0.0 * v.z,
            (-s * v.x + c * v.y) + 0.0 * v.z,
            (0.0 * v.x + 0.0 * v.y) + 1.0 * v.z};
}

inline Vec3 rot_global(const Vec3& v, const Vec3& no) {
    Vec3 n = normalized(no);
    double c = n.x, s = n.y;
    // R = [[c, -s, 0], [s, c, 0], [0, 0, 1]]
    return {(c * v.x + -s * v.y) + 0.0 * v.z,
            (s * v.x + c * v.y) + 0.0 * v.z,
            (0.0 * v.x + 0.0 * v.y) + 1.0 * v.z};
}

inline PState in_local(const PState& z, const Vec3& n) {
    PState r = z;
    r.v = rot_local(z.v, n);
    return r;
}

inline void flux_to_global(Flux& f, const Vec3& n) { f.m = rot_global(f.m, n); }

// ---------------------------------------------------------------- fluxes
thread_local int g_numeric_error = 0;

// HLLC::calc_flux, smf (math/cfd/flux/hllc.cpp:23-83).
Flux hllc(const PState& zL, const PState& zR, const Eos& eos) {
    double c_L = eos.sound_speed_rP(zL.rho, zL.p);
    double c_R = eos.sound_speed_rP(zR.rho, zR.p);

    double S_L = smin(zL.v.x - c_L, zR.v.x - c_R);
    double S_R = smax(zL.v.x + c_L, zR.v.x + c_R);

    if (S_L >= 0.0) { return phys_flux(zL); }
    if (S_R <= 0.0) { return phys_flux(zR); }

    double a_L = zL.rho * (S_L - zL.v.x);
    double a_R = zR.rho * (S_R - zR.v.x);

    double S_C = (zL.p - zR.p + a_R * zR.v.x - a_L * zL.v.x) / (a_R - a_L);

    double rho_sL = a_L / (S_L - S_C);
    double rho_sR = a_R / (S_R - S_C);

    double E_sL = total_E(zL) + (S_C - zL.v.x) * (S_C + zL.p / a_L);
    double E_sR = total_E(zR) + (S_C - zR.v.x) * (S_C + zR.p / a_R);

    QState Q_sL{rho_sL, {rho_sL * S_C, rho_sL * zL.v.y, rho_sL * zL.v.z}, rho_sL * E_sL};
    QState Q_sR{rho_sR, {rho_sR * S_C, rho_sR * zR.v.y, rho_sR * zR.v.z}, rho_sR * E_sR};

    QState Q_L = to_qstate(zL);
    QState Q_R = to_qstate(zR);
    Flux F_L = phys_flux(zL);
    Flux F_R = phys_flux(zR);

    double aSC = std::abs(S_C);
    Flux F;
    for (int k = 0; k < 5; ++k) {
        double t = F_L.arr()[k] + F_R.arr()[k];
        t = t + S_L * (Q_sL.arr()[k] - Q_L.arr()[k]);
        t = t + aSC * (Q_sL.arr()[k] - Q_sR.arr()[k]);
        t = t + S_R * (Q_sR.arr()[k] - Q_R.arr()[k]);
        F.arr()[k] = 0.5 * t;
        if (std::isnan(F.arr()[k])) { g_numeric_error = 1; }   // hllc.cpp:72-80 throws
    }
    return F;
}

// HLL::calc_flux, smf (math/cfd/flux/hll.cpp:23-57).
Flux hll(const PState& zL, const PState& zR, const Eos& eos) {
    double u_L = zL.v.x, u_R = zR.v.x;
    double c_L = eos.sound_speed_rP(zL.rho, zL.p);
    double c_R = eos.sound_speed_rP(zR.rho, zR.p);

    double S_L = smin(u_L - c_L, smin(u_R - c_R, 0.0));
    double S_R = smax(u_L + c_L, smax(u_R + c_R, 0.0));

    QState Q_L = to_qstate(zL), Q_R = to_qstate(zR);
    Flux F_L = phys_flux(zL), F_R = phys_flux(zR);

    double SLSR = S_L * S_R;
    double den = S_R - S_L;
    Flux F;
    for (int k = 0; k < 5; ++k) {
        double t = S_R * F_L.arr()[k] - S_L * F_R.arr()[k] + SLSR * (Q_R.arr()[k] - Q_L.arr()[k]);
        F.arr()[k] = t / den;
        if (std::isnan(F.arr()[k])) { g_numeric_error = 1; }   // hll.cpp:46-54 throws
    }
    return F;
}

// Rusanov::calc_flux, smf (math/cfd/flux/rusanov.cpp:13-40).  std::min of an
// initializer list keeps the first smallest element.
Flux rusanov(const PState& zL, const PState& zR, const Eos& eos) {
    double u_L = zL.v.x, u_R = zR.v.x;
    double c_L = eos.sound_speed_rP(zL.rho, zL.p);
    double c_R = eos.sound_speed_rP(zR.rho, zR.p);

    double S_L = u_L - c_L;
    if (u_R - c_R < S_L) S_L = u_R - c_R;
    if (0.0 < S_L) S_L = 0.0;
    double S_R = u_L + c_L;
    if (S_R < u_R + c_R) S_R = u_R + c_R;
    if (S_R < 0.0) S_R = 0.0;

    double S = smax(std::abs(S_L), S_R);

    QState Q_L = to_qstate(zL), Q_R = to_qstate(zR);
    Flux F_L = phys_flux(zL), F_R = phys_flux(zR);
    double hS = 0.5 * S;
    Flux F;
    for (int k = 0; k < 5; ++k) {
        F.arr()[k] = 0.5 * (F_L.arr()[k] + F_R.arr()[k]) + hS * (Q_L.arr()[k] - Q_R.arr()[k]);
    }
    return F;
}

// HLLC_M::calc_flux (math/cfd/flux/hllc.cpp:491-572): HLLC with the
// tangential velocities averaged across the contact and the matching
// energy correction; std::pow(x, 2) is x * x bit for bit.
Flux hllc_m(const PState& zL, const PState& zR, const Eos& eos) {
    double c_L = eos.sound_speed_rP(zL.rho, zL.p);
    double c_R = eos.sound_speed_rP(zR.rho, zR.p);
    double S_L = smin(zL.v.x - c_L, zR.v.x - c_R);
    double S_R = smax(zL.v.x + c_L, zR.v.x + c_R);
    if (S_L >= 0.0) { return phys_flux(zL); }
    if (S_R <= 0.0) { return phys_flux(zR); }
    double a_L = zL.rho * (S_L - zL.v.x);
    double a_R = zR.rho * (S_R - zR.v.x);
    double S_C = (zL.p - zR.p + a_R * zR.v.x - a_L * zL.v.x) / (a_R - a_L);
    double V_C = (a_R * zR.v.y - a_L * zL.v.y) / (a_R - a_L);
    double W_C = (a_R * zR.v.z - a_L * zL.v.z) / (a_R - a_L);
    double rho_sL = a_L / (S_L - S_C);
    double rho_sR = a_R / (S_R - S_C);
    double E_sL = total_E(zL) + (S_C - zL.v.x) * (S_C + zL.p / a_L);
    double E_sR = total_E(zR) + (S_C - zR.v.x) * (S_C + zR.p / a_R);
    double v_L2 = zL.v.y * zL.v.y + zL.v.z * zL.v.z;
    double v_R2 = zR.v.y * zR.v.y + zR.v.z * zR.v.z;
    E_sL += 0.5 * ((a_R * v_R2 - a_L * v_L2) / (a_R - a_L) - v_L2);
    E_sR += 0.5 * ((a_R * v_R2 - a_L * v_L2) / (a_R - a_L) - v_R2);
    QState Q_sL{rho_sL, {rho_sL * S_C, rho_sL * V_C, rho_sL * W_C}, rho_sL * E_sL};
    QState Q_sR{rho_sR, {rho_sR * S_C, rho_sR * V_C, rho_sR * W_C}, rho_sR * E_sR};
    QState Q_L = to_qstate(zL), Q_R = to_qstate(zR);
    Flux F_L = phys_flux(zL), F_R = phys_flux(zR);
    double aSC = std::abs(S_C);
    Flux F;
    for (int k = 0; k < 5; ++k) {
        double t = F_L.arr()[k] + F_R.arr()[k];
        t = t + S_L * (Q_sL.arr()[k] - Q_L.arr()[k]);
        t = t + aSC * (Q_sL.arr()[k] - Q_sR.arr()[k]);
        t = t + S_R * (Q_sR.arr()[k] - Q_R.arr()[k]);
        F.arr()[k] = 0.5 * t;
        if (std::isnan(F.arr()[k])) { g_numeric_error = 1; }   // hllc.cpp:558-568 throws
    }
    return F;
}

// HLLC_LM::calc_flux (math/cfd/flux/hllc.cpp:427-488): HLLC whose outer
// wave terms are scaled by phi = sin(pi/2 * Ma / 0.1) below Ma = 0.1 (1
// above), Ma the larger |v_x| / c of the two sides.
Flux hllc_lm(const PState& zL, const PState& zR, const Eos& eos) {
    double c_L = eos.sound_speed_rP(zL.rho, zL.p);
    double c_R = eos.sound_speed_rP(zR.rho, zR.p);
    double S_L = smin(zL.v.x - c_L, zR.v.x - c_R);
    double S_R = smax(zL.v.x + c_L, zR.v.x + c_R);
    if (S_L >= 0.0) { return phys_flux(zL); }
    if (S_R <= 0.0) { return phys_flux(zR); }
    double a_L = zL.rho * (S_L - zL.v.x);
    double a_R = zR.rho * (S_R - zR.v.x);
    double S_C = (zL.p - zR.p + a_R * zR.v.x - a_L * zL.v.x) / (a_R - a_L);
    double rho_sL = a_L / (S_L - S_C);
    double rho_sR = a_R / (S_R - S_C);
    double E_sL = total_E(zL) + (S_C - zL.v.x) * (S_C + zL.p / a_L);
    double E_sR = total_E(zR) + (S_C - zR.v.x) * (S_C + zR.p / a_R);
    QState Q_sL{rho_sL, {rho_sL * S_C, rho_sL * zL.v.y, rho_sL * zL.v.z}, rho_sL * E_sL};
    QState Q_sR{rho_sR, {rho_sR * S_C, rho_sR * zR.v.y, rho_sR * zR.v.z}, rho_sR * E_sR};
    QState Q_L = to_qstate(zL), Q_R = to_qstate(zR);
    Flux F_L = phys_flux(zL), F_R = phys_flux(zR);
    const double Ma_limit = 0.1;
    double Ma_local = smax(std::abs(zL.v.x / c_L), std::abs(zR.v.x / c_R));
    double xi = Ma_local / Ma_limit;
    double phi = xi > 1.0 ? 1.0 : std::sin(M_PI_2 * xi);
    double aSC = std::abs(S_C);
    Flux F;
    for (int k = 0; k < 5; ++k) {
        double t = F_L.arr()[k] + F_R.arr()[k];
        t = t + phi * S_L * (Q_sL.arr()[k] - Q_L.arr()[k]);
        t = t + aSC * (Q_sL.arr()[k] - Q_sR.arr()[k]);
        t = t + phi * S_R * (Q_sR.arr()[k] - Q_R.arr()[k]);
        F.arr()[k] = 0.5 * t;
        if (std::isnan(F.arr()[k])) { g_numeric_error = 1; }   // hllc.cpp:474-484 throws
    }
    return F;
}

// Arithmetic face state of the CIR fluxes (cir.cpp:22-27): the mean of the
// two primitive vectors with the energy made consistent with (rho, p).
inline PState cir_face_state(const PState& zL, const PState& zR, const Eos& eos) {
    PState zE;
    for (int k = 0; k < 6; ++k) zE.arr()[k] = 0.5 * (zL.arr()[k] + zR.arr()[k]);
    zE.e = eos.energy_rP(zE.rho, zE.p);
    return zE;
}

// CIR1::calc_flux (math/cfd/flux/cir.cpp:16-111): F = (F_L + F_R) / 2 +
// |A| (q_L - q_R) / 2 with |A| = R |Lambda| L, the eigen-decomposition of
// the flux Jacobian at the arithmetic face state.  The reference forms the
// products with Eigen (OmR * abs_L * OmL, then abs_A * dq); this
// restatement fixes the order: |A|(i,j) = sum over k ascending of
// (R(i,k) |lambda_k|) L(k,j), (|A| dq)_i summed over j ascending.
Flux cir1(const PState& zL, const PState& zR, const Eos& eos) {
    const PState zE = cir_face_state(zL, zR, eos);
    const double dPdE = (eos.gamma - 1.0) * zE.rho;   // pressure_re(..., deriv).dE
    const double rho = zE.rho, u = zE.v.x, v = zE.v.y, w = zE.v.z, p = zE.p, eps = zE.e;
    const double u2 = u * u, v2 = v * v, w2 = w * w;
    const double q2 = u2 + v2 + w2;
    const double E = rho * (eps + 0.5 * q2);
    const double c = eos.sound_speed_rP(rho, p);
    const double c2 = c * c;
    const double h = (E + p) / rho;
    const double b = dPdE / rho;
    const double T = b * (q2 - h) + c2;
    const double lam[5] = {std::abs(u - c), std::abs(u), std::abs(u), std::abs(u), std::abs(u + c)};
    const double R[5][5] = {{1, 0, 0, 1, 1},
                            {u - c, 0, 0, u, u + c},
                            {v, 1, 0, v, v},
                            {w, 0, 1, w, w},
                            {h - u * c, v, w, h - c2 / b, h + u * c}};
    double L[5][5] = {{T + u * c, -b * u - c, -b * v, -b * w, b},
                      {-2 * v * c2, 0, 2 * c2, 0, 0},
                      {-2 * w * c2, 0, 0, 2 * c2, 0},
                      {2 * b * (h - q2), 2 * b * u, 2 * b * v, 2 * b * w, -2 * b},
                      {T - u * c, -b * u + c, -b * v, -b * w, b}};
    const double s2 = 2.0 * c2;
    for (auto& row : L)
        for (double& x : row) x = x / s2;
    const QState qL = to_qstate(zL), qR = to_qstate(zR);
    const Flux fL = phys_flux(zL), fR = phys_flux(zR);
    double dq[5];
    for (int k = 0; k < 5; ++k) dq[k] = qL.arr()[k] - qR.arr()[k];
    Flux F;
    for (int i = 0; i < 5; ++i) {
        double t = 0.0;
        for (int j = 0; j < 5; ++j) {
            double a = 0.0;
            for (int k = 0; k < 5; ++k) a = a + (R[i][k] * lam[k]) * L[k][j];
            t = t + a * dq[j];
        }
        F.arr()[i] = 0.5 * (fL.arr()[i] + fR.arr()[i]) + 0.5 * t;
    }
    return F;
}

// CIR2::calc_flux (math/cfd/flux/cir.cpp:117-216): the same with the
// pressure as a sixth variable, which needs no derivative of the EOS: a
// 5 x 6 R, a 6 x 6 L, q extended by p.  The reference checks at run time
// that R Lambda L is the extended Jacobian and R L the identity (1e-10
// relative) and throws otherwise; both hold to round-off for every finite
// state, and the restatement does not repeat the check.
Flux cir2(const PState& zL, const PState& zR, const Eos& eos) {
    const PState zE = cir_face_state(zL, zR, eos);
    const double rho = zE.rho, u = zE.v.x, v = zE.v.y, w = zE.v.z, p = zE.p, eps = zE.e;
    const double u2 = u * u, v2 = v * v, w2 = w * w;
    const double q2 = u2 + v2 + w2;
    const double E = rho * (eps + 0.5 * q2);
    const double c = eos.sound_speed_rP(rho, p);
    const double c2 = c * c;
    const double h = (E + p) / rho;
    const double lam[6] = {std::abs(u - c), std::abs(u), std::abs(u), std::abs(u), std::abs(u + c), std::abs(u)};
    const double R[5][6] = {{1, 0, 0, 0, 1, 1},
                            {u - c, 0, 0, 0, u + c, u},
                            {v, 1, 0, 0, v, 0},
                            {w, 0, 1, 0, w, 0},
                            {h - u * c, 0, 0, 1, h + u * c, 0}};
    const double L[6][6] = {{u / (2 * c), -1 / (2 * c), 0, 0, 0, 1 / (2 * c2)},
                            {0, 0, 1, 0, 0, -v / c2},
                            {0, 0, 0, 1, 0, -w / c2},
                            {u2, -u, 0, 0, 1, -h / c2},
                            {-u / (2 * c), 1 / (2 * c), 0, 0, 0, 1 / (2 * c2)},
                            {1, 0, 0, 0, 0, -1 / c2}};
    const QState qL = to_qstate(zL), qR = to_qstate(zR);
    const Flux fL = phys_flux(zL), fR = phys_flux(zR);
    double dq[6];
    for (int k = 0; k < 5; ++k) dq[k] = qL.arr()[k] - qR.arr()[k];
    dq[5] = zL.p - zR.p;
    Flux F;
    for (int i = 0; i < 5; ++i) {
        double t = 0.0;
        for (int j = 0; j < 6; ++j) {
            double a = 0.0;
            for (int k = 0; k < 6; ++k) a = a + (R[i][k] * lam[k]) * L[k][j];
            t = t + a * dq[j];
        }
        F.arr()[i] = 0.5 * (fL.arr()[i] + fR.arr()[i]) + 0.5 * t;
    }
    return F;
}

// ---------------------------------------------------------------- exact Riemann
// RiemannSolver (math/solver/riemann.cpp:12-460), stiffened-gas two-term EOS.
inline double sqr(double x) { return x * x; }
inline double eos_sound(double rho, double p, double g, double p0) { return std::sqrt(g * (p + p0) / rho); }
inline double eos_density(double p, double c, double g, double p0) { return g * (p + p0) / sqr(c); }

inline double p_init(double rL, double uL, double pL, double cL, double p0L,
                     double rR, double uR, double pR, double cR, double p0R) {
    double rcL = rL * cL;
    double rcR = rR * cR;
    double P = (pL * rcR + pR * rcL + (uL - uR) * rcL * rcR) / (rcL + rcR);
    double Pmin = -smin(p0L, p0R) + 1.0e-2 * smin(pL + p0L, pR + p0R);
    return smax(Pmin, P);
}

inline double coeff_AK(double rK, double gK) { return 2.0 / ((gK + 1) * rK); }
inline double coeff_BK(double pK, double gK, double p0K) {
    return (gK - 1.0) * pK / (gK + 1.0) + 2.0 * gK * p0K / (gK + 1.0);
}
inline double coeff_GK(double g) { return (g - 1.0) / (2.0 * g); }

inline double func_fK(double P, double rK, double pK, double cK, double gK, double p0K,
                      double AK, double BK, double GK) {
    (void)rK;
    if (P >= pK) {
        return (P - pK) * std::sqrt(AK / (P + BK));
    }
    return 2.0 * cK / (gK - 1.0) * (std::pow((P + p0K) / (pK + p0K), GK) - 1.0);
}

inline double deriv_fK(double P, double rK, double pK, double cK, double gK, double p0K,
                       double AK, double BK, double GK) {
    (void)gK;
    if (P >= pK) {
        return (1.0 - (P - pK) / (2.0 * (P + BK))) * std::sqrt(AK / (P + BK));
    }
    return std::pow((P + p0K) / (pK + p0K), GK - 1.0) / (rK * cK);
}

struct SolPU { double P, U; bool conv; };

// contact_p (riemann.cpp:118-234).
SolPU contact_p(double rL, double uL, double pL, double cL, double gL, double p0L,
                double rR, double uR, double pR, double cR, double gR, double p0R) {
    double AL = coeff_AK(rL, gL), AR = coeff_AK(rR, gR);
    double BL = coeff_BK(pL, gL, p0L), BR = coeff_BK(pR, gR, p0R);
    double GL = coeff_GK(gL), GR = coeff_GK(gR);

    double p_dno = smax(-p0L, -p0R);
    double p_min = smin(pL, pR);
    double p_max = smax(pL, pR);

    double f_dno = func_fK(p_dno, rL, pL, cL, gL, p0L, AL, BL, GL) +
                   func_fK(p_dno, rR, pR, cR, gR, p0R, AR, BR, GR) + uR - uL;
    if (f_dno > 0.0) {
        return {p_dno, 0.0, true};   // vacuum
    }

    double P = 0.5 * (pL + pR);
    double f_min = func_fK(p_min, rL, pL, cL, gL, p0L, AL, BL, GL) +
                   func_fK(p_min, rR, pR, cR, gR, p0R, AR, BR, GR) + uR - uL;
    if (f_min > 0.0) {
        P = p_init(rL, uL, pL, cL, p0L, rR, uR, pR, cR, p0R);
    } else {
        double f_max = func_fK(p_max, rL, pL, cL, gL, p0L, AL, BL, GL) +
                       func_fK(p_max, rR, pR, cR, gR, p0R, AR, BR, GR) + uR - uL;
        (void)f_max;   // every branch uses the acoustic initial guess
        P = p_init(rL, uL, pL, cL, p0L, rR, uR, pR, cR, p0R);
    }

    const int max_iterations = 50;
    double DP1 = 1.0e-5 * std::abs(pL - pR);
    double DP2 = 1.0e-12 * std::abs(smax(pL + p0L, pR + p0R));
    double DP = smax(DP1, DP2);

    double fL = 0, fR = 0, dfL, dfR;
    int counter = 0;
    double Pn = P;
    for (; counter < max_iterations; ++counter) {
        fL = func_fK(P, rL, pL, cL, gL, p0L, AL, BL, GL);
        fR = func_fK(P, rR, pR, cR, gR, p0R, AR, BR, GR);
        dfL = deriv_fK(P, rL, pL, cL, gL, p0L, AL, BL, GL);
        dfR = deriv_fK(P, rR, pR, cR, gR, p0R, AR, BR, GR);

        Pn = P - (fL + fR + uR - uL) / (dfL + dfR);
        if (Pn < p_dno) {
            Pn = (p_dno + P) / 2;
        }
        if (std::abs(Pn - P) < DP) {
            if (counter > 3) break;
        }
        P = Pn;
    }
    bool conv = counter < max_iterations;
    if (!conv && std::abs((P - Pn) / (P + Pn)) > 0.02) {
        g_numeric_error = 1;   // riemann.cpp:222-229 throws
    }
    double U = 0.5 * (uL + uR + fR - fL);
    return {P, U, conv};
}

inline double shock_wave_speed(double rK, double uK, double pK, double U, double P) {
    return uK + (P - pK) / (rK * (U - uK));
}
inline double shock_wave_density(double rK, double uK, double pK, double U, double P) {
    return rK * (P - pK) / ((P - pK) - rK * sqr(U - uK));
}
inline double rarefaction_sound_L(double U, double uL, double cL, double gL) { return cL + 0.5 * (gL - 1.0) * (uL - U); }
inline double rarefaction_sound_R(double U, double uR, double cR, double gR) { return cR - 0.5 * (gR - 1.0) * (uR - U); }
inline double density_lfan(double rL, double uL, double cL, double gL, double xi) {
    return rL * std::pow(2.0 / (gL + 1.0) + ((gL - 1.0) * (uL - xi)) / ((gL + 1.0) * cL), 2.0 / (gL - 1.0));
}
inline double density_rfan(double rR, double uR, double cR, double gR, double xi) {
    return rR * std::pow(2.0 / (gR + 1.0) - ((gR - 1.0) * (uR - xi)) / ((gR + 1.0) * cR), 2.0 / (gR - 1.0));
}
inline double velocity_lfan(double uL, double cL, double gL, double xi) { return ((gL - 1.0) * uL + 2.0 * (cL + xi)) / (gL + 1.0); }
inline double velocity_rfan(double uR, double cR, double gR, double xi) { return ((gR - 1.0) * uR + 2.0 * (xi - cR)) / (gR + 1.0); }
inline double pressure_lfan(double uL, double pL, double cL, double gL, double p0L, double xi) {
    return (pL + p0L) * std::pow(2.0 / (gL + 1.0) + ((gL - 1.0) * (uL - xi)) / ((gL + 1.0) * cL), 2.0 * gL / (gL - 1.0)) - p0L;
}
inline double pressure_rfan(double uR, double pR, double cR, double gR, double p0R, double xi) {
    return (pR + p0R) * std::pow(2.0 / (gR + 1.0) - ((gR - 1.0) * (uR - xi)) / ((gR + 1.0) * cR), 2.0 * gR / (gR - 1.0)) - p0R;
}

// Field order follows the Solution(rho, Uf, Pf, U, P, conv) constructor
// (riemann.h:35-36); every return below lists U before P.
struct RSol { double rho, Uf, Pf, U, P; bool conv; };

// RiemannSolver::solve (riemann.cpp:366-460): the face state at xi = 0.
RSol riemann_solve(double rL, double uL, double pL, double gL, double p0L,
                   double rR, double uR, double pR, double gR, double p0R) {
    double cL = eos_sound(rL, pL, gL, p0L);
    double cR = eos_sound(rR, pR, gR, p0R);
    SolPU PU = contact_p(rL, uL, pL, cL, gL, p0L, rR, uR, pR, cR, gR, p0R);
    double P = PU.P, U = PU.U;
    bool conv = PU.conv;
    if (U > 0.0) {
        if (pL < P) {
            double D = shock_wave_speed(rL, uL, pL, U, P);
            if (D > 0.0) return {rL, uL, pL, U, P, conv};
            return {shock_wave_density(rL, uL, pL, U, P), U, P, U, P, conv};
        }
        double DL1 = uL - cL;
        if (DL1 > 0.0) return {rL, uL, pL, U, P, conv};
        double cl = rarefaction_sound_L(U, uL, cL, gL);
        double DL2 = U - cl;
        if (DL2 < 0.0) return {eos_density(P, cl, gL, p0L), U, P, U, P, conv};
        return {density_lfan(rL, uL, cL, gL, 0.0), velocity_lfan(uL, cL, gL, 0.0),
                pressure_lfan(uL, pL, cL, gL, p0L, 0.0), U, P, conv};
    }
    if (P > pR) {
        double D = shock_wave_speed(rR, uR, pR, U, P);
        if (D < 0.0) return {rR, uR, pR, U, P, conv};
        return {shock_wave_density(rR, uR, pR, U, P), U, P, U, P, conv};
    }
    double DR2 = uR + cR;
    if (DR2 < 0.0) return {rR, uR, pR, U, P, conv};
    double cr = rarefaction_sound_R(U, uR, cR, gR);
    double DR1 = U + cr;
    if (DR1 > 0.0) return {eos_density(P, cr, gR, p0R), U, P, U, P, conv};
    return {density_rfan(rR, uR, cR, gR, 0.0), velocity_rfan(uR, cR, gR, 0.0),
            pressure_rfan(uR, pR, cR, gR, p0R, 0.0), U, P, conv};
}

// Godunov::calc_flux, smf (math/cfd/flux/godunov.cpp:16-32).
Flux godunov(const PState& zL, const PState& zR, const Eos& eos) {
    RSol sol = riemann_solve(zL.rho, zL.v.x, zL.p, eos.gamma, eos.P0,
                             zR.rho, zR.v.x, zR.p, eos.gamma, eos.P0);
    Vec3 v{sol.Uf, sol.U > 0.0 ? zL.v.y : zR.v.y, sol.U > 0.0 ? zL.v.z : zR.v.z};
    PState z{sol.rho, v, sol.Pf, eos.energy_rP(sol.rho, sol.Pf)};
    return phys_flux(z);
}

Flux num_flux(int kind, const PState& zL, const PState& zR, const Eos& eos) {
    switch (kind) {
        case ORC_GODUNOV: return godunov(zL, zR, eos);
        case ORC_RUSANOV: return rusanov(zL, zR, eos);
        case ORC_HLL:     return hll(zL, zR, eos);
        case ORC_HLLC_M:  return hllc_m(zL, zR, eos);
        case ORC_HLLC_LM: return hllc_lm(zL, zR, eos);
        case ORC_CIR1:    return cir1(zL, zR, eos);
        case ORC_CIR2:    return cir2(zL, zR, eos);
        default:          return hllc(zL, zR, eos);
    }
}

// ---------------------------------------------------------------- limiters
// Binary limiter functions func2(a, b) (math/cfd/limiter.cxx:15-153).
double lim_unary(int kind, double r) {
    switch (kind) {
        case ORC_LIM_NONE: return 0.0;
        case ORC_LIM_UNIT: return 1.0;
        case ORC_LIM_MINMOD:   // :35-38
            if (std::isnan(r)) return 0.0;
            return smax(0.0, smin(r, 1.0));
        case ORC_LIM_MC:       // :45-48
            if (std::isnan(r)) return 0.0;
            return smax(0.0, smin(2 * r, smin(0.5 * (1.0 + r), 2.0)));
        case ORC_LIM_VANLEER:  // :54-60
            if (r <= 0.0) return 0.0;
            if (std::isnan(r)) return 0.0;
            if (std::isinf(r)) return 2.0;
            return 2.0 * r / (r + 1.0);
        case ORC_LIM_VANALBADA1:   // :69-75
            if (r <= 0.0) return 0.0;
            if (std::isnan(r)) return 0.0;
            if (std::isinf(r)) return 1.0;
            return r * (r + 1.0) / (sqr(r) + 1.0);
        case ORC_LIM_VANALBADA2:   // :84-90
            if (r <= 0.0) return 0.0;
            if (std::isnan(r)) return 0.0;
            if (std::isinf(r)) return 0.0;
            return 2.0 * r / (sqr(r) + 1.0);
        case ORC_LIM_CHARM:    // :98-101
            if (std::isnan(r)) return 0.0;
            return r > 0.0 ? (3.0 + 1.0 / r) / sqr(r + 1.0 / r) : 0.0;
        case ORC_LIM_HCUS:     // :108-111
            if (std::isnan(r)) return 0.0;
            return r > 0.0 ? 3.0 / (1.0 + 2.0 / r) : 0.0;
        case ORC_LIM_HQUICK:   // :118-121
            if (std::isnan(r)) return 0.0;
            return r > 0.0 ? 4.0 / (1.0 + 3.0 / r) : 0.0;
        case ORC_LIM_KOREN:    // :128-131
            if (std::isnan(r)) return 0.0;
            return smax(0.0, smin(2 * r, smin(2.0, (1 + 2 * r) / 3.0)));
        case ORC_LIM_SMART:    // :137-140
            if (std::isnan(r)) return 0.0;
            return smax(0.0, smin(2.0 * r, smin(0.25 + 0.75 * r, 4.0)));
        case ORC_LIM_SUPERBEE: // :146-149
            if (std::isnan(r)) return 0.0;
            return smax(0.0, smin(smin(2 * r, 1.0), smin(r, 2.0)));
    }
    return 0.0;
}

double lim_binary(int kind, double a, double b) {
    switch (kind) {
        case ORC_LIM_NONE: return 0.0;
        case ORC_LIM_UNIT: return 1.0;
        case ORC_LIM_MINMOD:   // :40-43
            if (a * b <= 0.0) return 0.0;
            return smin(a / b, 1.0);
        case ORC_LIM_VANLEER:  // :62-67
            if (a * b < 0.0) return 0.0;
            if (a == 0.0) return 0.0;
            return 2 * std::abs(a) / (std::abs(a) + std::abs(b));
        case ORC_LIM_VANALBADA1:   // :77-82
            if (a * b <= 0.0) return 0.0;
            if (b == 0.0) return 1.0;
            return a * (a + b) / (sqr(a) + sqr(b));
        case ORC_LIM_VANALBADA2: { // :92-96
            if (a * b <= 0.0) return 0.0;
            double eps = 10e-6;
            return (2.0 * a * b + eps) / (sqr(a) + sqr(b) + eps);
        }
        case ORC_LIM_CHARM:    // :103-106
            if (a * b <= 0.0) return 0.0;
            return a * (3 * a + b) / sqr(a + b);
        case ORC_LIM_HCUS:     // :113-116
            if (a * b <= 0.0) return 0.0;
            return (3.0 * a) / (a + 2.0 * b);
        case ORC_LIM_HQUICK:   // :123-126
            if (a * b <= 0.0) return 0.0;
            return (4.0 * a) / (a + 3.0 * b);
        default:               // MC, Koren, smart, superbee: func1(a / b)
            return lim_unary(kind, a / b);
    }
}

}  // namespace

// =============================================================== mesh
struct orc_mesh {
    int n = 0;
    // the structured (i, j) grid of a Rectangle, cell = gny i + j (EuMesh::nx,
    // ny and operator()(i, j), eu_mesh.cpp:826-830); 0 on other meshes
    int gnx = 0, gny = 0;
    // cells (amr_cells.h:74-104); idx < 0 marks an undefined cell during AMR
    std::vector<double> cx, cy, vol;
    std::vector<int> level, flag, b_idx, z_idx, next, idx;
    std::vector<double> vx, vy;   // 9 per cell, SqQuad order (quad.h:186-191)
    // face slots, 8 per cell (amr_faces.h:83-101)
    std::vector<int> bnd, adj;
    std::vector<double> fnx, fny, fcx, fcy, area;
    // axisymmetric geometry (amr_cells.h volume_alt, amr_faces.h area_alt):
    // y-weighted volume and face areas, with y-shifted centres
    bool axial = false;
    std::vector<double> vol_alt, area_alt;
};

namespace {

inline bool slot_is_boundary(int b) {   // amr_faces.h:157-161
    return b != ORC_INNER && b != ORC_PERIODIC && b != ORC_UNDEFINED;
}

// EuFace::symm_point (amr_faces.h:197-199).
inline Vec3 symm_point(const orc_mesh& m, int s, const Vec3& p) {
    Vec3 fc{m.fcx[s], m.fcy[s], 0.0};
    Vec3 n{m.fnx[s], m.fny[s], 0.0};
    double d = 2.0 * dot(fc - p, n);
    return p + scale(d, n);
}

// boundary_value (math/solver/sm_fluid.cpp:69-79).
inline PState boundary_value(const PState& zc, const Vec3& normal, int flag) {
    if (flag != ORC_WALL) return zc;
    PState zn = zc;
    double d = dot(zc.v, normal);
    Vec3 Vn = scale(d, normal);
    zn.v = zc.v - scale(2.0, Vn);
    return zn;
}

}  // namespace

extern "C" {

// AmrCells(const Rectangle&) / Rectangle::initialize (mesh/euler/amr_cells_gen.cpp:128-281,
// geom/generator/rectangle.cpp:465-603); axial: Rectangle::set_axial
// (rectangle.cpp:97-103, a bottom on the axis becomes a wall) and the
// alternative volume, face areas and shifted centres (rectangle.cpp:590-602).
orc_mesh* orc_mesh_rectangle_ax(int nx, int ny, double xmin, double xmax,
                                double ymin, double ymax, const int* bounds_in, int axial) {
    int bounds[4] = {bounds_in[0], bounds_in[1], bounds_in[2], bounds_in[3]};
    if (axial && ymin == 0.0) bounds[2] = ORC_WALL;
    orc_mesh* m = new orc_mesh();
    m->axial = axial != 0;
    if (m->axial) {
        m->vol_alt.resize((size_t)nx * ny);
        m->area_alt.assign(8 * (size_t)nx * ny, 0.0);
    }
    int n = nx * ny;
    m->n = n;
    m->gnx = nx;
    m->gny = ny;
    m->cx.resize(n); m->cy.resize(n); m->vol.resize(n);
    m->level.assign(n, 0); m->flag.assign(n, 0); m->b_idx.resize(n); m->z_idx.assign(n, 0);
    m->next.resize(n); m->idx.resize(n);
    for (int c = 0; c < n; ++c) { m->next[c] = c; m->idx[c] = c; }
    m->vx.resize(9 * (size_t)n); m->vy.resize(9 * (size_t)n);
    size_t ns = 8 * (size_t)n;
    m->bnd.assign(ns, ORC_UNDEFINED); m->adj.assign(ns, -1);
    m->fnx.assign(ns, 0.0); m->fny.assign(ns, 0.0);
    m->fcx.assign(ns, 0.0); m->fcy.assign(ns, 0.0); m->area.assign(ns, 0.0);

    const double hx = (xmax - xmin) / nx;
    const double hy = (ymax - ymin) / ny;
    auto vx_of = [&](int i) { return xmin + ((xmax - xmin) * i) / nx; };
    auto vy_of = [&](int j) { return ymin + ((ymax - ymin) * j) / ny; };
    const bool xp = bounds[0] == ORC_PERIODIC && bounds[1] == ORC_PERIODIC;
    const bool yp = bounds[2] == ORC_PERIODIC && bounds[3] == ORC_PERIODIC;

    for (int ic = 0; ic < n; ++ic) {
        int i = ic / ny, j = ic % ny;
        m->b_idx[ic] = ic;
        // SqQuad(v00, v01, v10, v11) (geom/primitives/quad.cpp:429-436)
        Vec3 v00{vx_of(i), vy_of(j), 0.0}, v01{vx_of(i + 1), vy_of(j), 0.0};
        Vec3 v10{vx_of(i), vy_of(j + 1), 0.0}, v11{vx_of(i + 1), vy_of(j + 1), 0.0};
        Vec3 q[9] = {
            v00, scale(0.5, v00 + v01), v01,
            scale(0.5, v00 + v10), scale(0.25, ((v00 + v01) + v10) + v11), scale(0.5, v01 + v11),
            v10, scale(0.5, v10 + v11), v11};
        for (int k = 0; k < 9; ++k) { m->vx[9 * (size_t)ic + k] = q[k].x; m->vy[9 * (size_t)ic + k] = q[k].y; }
        m->cx[ic] = q[4].x; m->cy[ic] = q[4].y;
        m->vol[ic] = hx * hy;

        size_t f = 8 * (size_t)ic;
        m->bnd[f + 0] = i > 0 ? ORC_INNER : bounds[0];
        m->bnd[f + 1] = i < nx - 1 ? ORC_INNER : bounds[1];
        m->bnd[f + 2] = j > 0 ? ORC_INNER : bounds[2];
        m->bnd[f + 3] = j < ny - 1 ? ORC_INNER : bounds[3];
        // neib_index (amr_cells_gen.cpp:167-184)
        m->adj[f + 0] = (i == 0 && !xp) ? ic : ((i - 1 + nx) % nx) * ny + j;
        m->adj[f + 1] = (i == nx - 1 && !xp) ? ic : ((i + 1 + nx) % nx) * ny + j;
        m->adj[f + 2] = (j == 0 && !yp) ? ic : i * ny + (j - 1 + ny) % ny;
        m->adj[f + 3] = (j == ny - 1 && !yp) ? ic : i * ny + (j + 1 + ny) % ny;
        m->fnx[f + 0] = -1.0; m->fny[f + 0] = 0.0;
        m->fnx[f + 1] = 1.0;  m->fny[f + 1] = 0.0;
        m->fnx[f + 2] = 0.0;  m->fny[f + 2] = -1.0;
        m->fnx[f + 3] = 0.0;  m->fny[f + 3] = 1.0;
        // face centres vs<-1,0>, vs<+1,0>, vs<0,-1>, vs<0,+1>
        m->fcx[f + 0] = q[3].x; m->fcy[f + 0] = q[3].y;
        m->fcx[f + 1] = q[5].x; m->fcy[f + 1] = q[5].y;
        m->fcx[f + 2] = q[1].x; m->fcy[f + 2] = q[1].y;
        m->fcx[f + 3] = q[7].x; m->fcy[f + 3] = q[7].y;
        m->area[f + 0] = hy; m->area[f + 1] = hy;
        m->area[f + 2] = hx; m->area[f + 3] = hx;
        if (m->axial) {
            m->vol_alt[ic] = hx * hy * q[4].y;
            m->area_alt[f + 0] = hy * q[3].y;
            m->area_alt[f + 1] = hy * q[5].y;
            m->area_alt[f + 2] = hx * q[1].y;
            m->area_alt[f + 3] = hx * q[7].y;
            m->cy[ic] += hy * hy / (12.0 * q[4].y);
            m->fcy[f + 0] += hy * hy / (12.0 * q[3].y);
            m->fcy[f + 1] += hy * hy / (12.0 * q[5].y);
        }
    }
    return m;
}

orc_mesh* orc_mesh_rectangle(int nx, int ny, double xmin, double xmax,
                             double ymin, double ymax, const int* bounds) {
    return orc_mesh_rectangle_ax(nx, ny, xmin, xmax, ymin, ymax, bounds, 0);
}

int orc_mesh_axial(const orc_mesh* m) { return m->axial ? 1 : 0; }

void orc_mesh_get_axial(const orc_mesh* m, double* vol_alt, double* area_alt) {
    if (!m->axial) return;
    std::memcpy(vol_alt, m->vol_alt.data(), (size_t)m->n * 8);
    std::memcpy(area_alt, m->area_alt.data(), 8 * (size_t)m->n * 8);
}

void orc_mesh_free(orc_mesh* m) { delete m; }
int orc_mesh_size(const orc_mesh* m) { return m->n; }

void orc_mesh_get_cells(const orc_mesh* m, double* cx, double* cy, double* vol,
                        int* level, int* flag, int* b_idx, int* z_idx) {
    size_t n = m->n;
    std::memcpy(cx, m->cx.data(), n * 8); std::memcpy(cy, m->cy.data(), n * 8);
    std::memcpy(vol, m->vol.data(), n * 8);
    std::memcpy(level, m->level.data(), n * 4); std::memcpy(flag, m->flag.data(), n * 4);
    std::memcpy(b_idx, m->b_idx.data(), n * 4); std::memcpy(z_idx, m->z_idx.data(), n * 4);
}

void orc_mesh_get_slots(const orc_mesh* m, int* boundary, int* adj, double* nx, double* ny,
                        double* fcx, double* fcy, double* area) {
    size_t ns = 8 * (size_t)m->n;
    std::memcpy(boundary, m->bnd.data(), ns * 4); std::memcpy(adj, m->adj.data(), ns * 4);
    std::memcpy(nx, m->fnx.data(), ns * 8); std::memcpy(ny, m->fny.data(), ns * 8);
    std::memcpy(fcx, m->fcx.data(), ns * 8); std::memcpy(fcy, m->fcy.data(), ns * 8);
    std::memcpy(area, m->area.data(), ns * 8);
}

void orc_mesh_set_slots(orc_mesh* m, const int* boundary, const int* adj, const double* nx, const double* ny,
                        const double* fcx, const double* fcy, const double* area) {
    size_t ns = 8 * (size_t)m->n;
    std::memcpy(m->bnd.data(), boundary, ns * 4); std::memcpy(m->adj.data(), adj, ns * 4);
    std::memcpy(m->fnx.data(), nx, ns * 8); std::memcpy(m->fny.data(), ny, ns * 8);
    std::memcpy(m->fcx.data(), fcx, ns * 8); std::memcpy(m->fcy.data(), fcy, ns * 8);
    std::memcpy(m->area.data(), area, ns * 8);
}

void orc_mesh_get_verts(const orc_mesh* m, double* vx, double* vy) {
    size_t nv = 9 * (size_t)m->n;
    std::memcpy(vx, m->vx.data(), nv * 8); std::memcpy(vy, m->vy.data(), nv * 8);
}

}  // extern "C"

namespace {

// AmrCells::incircle_diameter, adaptive 2D (mesh/euler/amr_cells.cpp:107-114).
inline double incircle(const orc_mesh& m, int c) {
    const double* X = &m.vx[9 * (size_t)c];
    const double* Y = &m.vy[9 * (size_t)c];
    // vs<i,j> = verts[3j + i + 4]
    Vec3 dxv{X[5] - X[3], Y[5] - Y[3], 0.0};
    Vec3 dyv{X[7] - X[1], Y[7] - Y[1], 0.0};
    return std::sqrt(smin(sqnorm(dxv), sqnorm(dyv)));
}

}  // namespace

extern "C" void orc_mesh_incircle(const orc_mesh* m, double* h) {
    for (int c = 0; c < m->n; ++c) h[c] = incircle(*m, c);
}

// =============================================================== solver
struct orc_smf {
    orc_mesh* m;
    Eos eos;
    int flux, limiter, acc;
    double cfl, dt, max_dt;
    // SmFluid::Parts (math/solver/sm_fluid.h:26-36)
    std::vector<PState> init, half, next, d_dx, d_dy, d_dz;
    std::vector<int> wait, wait_next;
    bool axial = false;   // SmFluid::set_axial (sm_fluid.cpp:45-47)
};

namespace {

struct Grad { PState x, y, z; };

inline void zero(PState& z) { std::memset(&z, 0, sizeof(z)); }

// gradient::LSM (Storable version, math/cfd/gradient.h:226-282) and
// gradient::limiting (gradient.h:412-474), as called by
// SmFluid::compute_grad (sm_fluid.cpp:116-125).
// raw (optional): the unlimited LSM gradient of rho, {d/dx, d/dy}.
void grad_cell(orc_smf& S, int c, Grad& out, double* raw = nullptr) {
    const orc_mesh& m = *S.m;
    const double w0 = 1.0;   // cell.dim() < 3
    const PState zc = S.init[c];
    const Vec3 cell_c{m.cx[c], m.cy[c], 0.0};

    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double Fx[6] = {0}, Fy[6] = {0}, Fz[6] = {0};

    struct FaceGeo { Vec3 dr, drn; double w; PState zn; };
    FaceGeo fg[8];
    int nf = 0;

    for (int s8 = 0; s8 < 8; ++s8) {
        size_t s = 8 * (size_t)c + s8;
        int b = m.bnd[s];
        if (b == ORC_UNDEFINED) continue;
        Vec3 normal{m.fnx[s], m.fny[s], 0.0};
        PState zn = !slot_is_boundary(b) ? S.init[m.adj[s]] : boundary_value(zc, normal, b);
        Vec3 nc = (b == ORC_INNER) ? Vec3{m.cx[m.adj[s]], m.cy[m.adj[s]], 0.0}
                                   : symm_point(m, (int)s, cell_c);
        Vec3 dr = nc - cell_c;
        Vec3 drn = scale(dot(dr, normal), normal);
        double w = w0 * m.area[s] / std::pow(norm(drn), 3);
        double wd[3] = {w * drn.x, w * drn.y, w * drn.z};
        double drv[3] = {dr.x, dr.y, dr.z};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] = A[i][j] + wd[i] * drv[j];
        for (int k = 0; k < 6; ++k) {
            double dF = zn.arr()[k] - zc.arr()[k];
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
    // Matrix3d::inverse, cofactor formula (Eigen InverseImpl.h, size 3).
    auto cof = [&](int i, int j) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1];
    };
    double c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    double det = (c00 * A[0][0] + c10 * A[1][0]) + c20 * A[2][0];
    double invdet = 1.0 / det;
    double B[3][3];
    B[0][0] = c00 * invdet; B[0][1] = c10 * invdet; B[0][2] = c20 * invdet;
    B[1][0] = cof(0, 1) * invdet; B[1][1] = cof(1, 1) * invdet; B[1][2] = cof(2, 1) * invdet;
    B[2][0] = cof(0, 2) * invdet; B[2][1] = cof(1, 2) * invdet; B[2][2] = cof(2, 2) * invdet;

    Grad g;
    for (int k = 0; k < 6; ++k) {
        g.x.arr()[k] = (B[0][0] * Fx[k] + B[0][1] * Fy[k]) + B[0][2] * Fz[k];
        g.y.arr()[k] = (B[1][0] * Fx[k] + B[1][1] * Fy[k]) + B[1][2] * Fz[k];
        g.z.arr()[k] = (B[2][0] * Fx[k] + B[2][1] * Fy[k]) + B[2][2] * Fz[k];
    }

    if (raw) {
        raw[0] = g.x.rho;
        raw[1] = g.y.rho;
    }

    // limiting: the same A (recomputed identically in the reference), new F.
    double Lx[6] = {0}, Ly[6] = {0}, Lz[6] = {0};
    for (int f = 0; f < nf; ++f) {
        const FaceGeo& q = fg[f];
        for (int k = 0; k < 6; ++k) {
            double dF_p = q.zn.arr()[k] - zc.arr()[k];
            double du_n = (g.x.arr()[k] * q.dr.x + g.y.arr()[k] * q.dr.y) + g.z.arr()[k] * q.dr.z;
            double dF_m = 2.0 * du_n - dF_p;
            double dF_lim = lim_binary(S.limiter, dF_m, dF_p) * dF_p;
            Lx[k] = Lx[k] + (q.w * dF_lim) * q.drn.x;
            Ly[k] = Ly[k] + (q.w * dF_lim) * q.drn.y;
            Lz[k] = Lz[k] + (q.w * dF_lim) * q.drn.z;
        }
    }
    for (int k = 0; k < 6; ++k) {
        out.x.arr()[k] = (B[0][0] * Lx[k] + B[0][1] * Ly[k]) + B[0][2] * Lz[k];
        out.y.arr()[k] = (B[1][0] * Lx[k] + B[1][1] * Ly[k]) + B[1][2] * Lz[k];
        out.z.arr()[k] = (B[2][0] * Lx[k] + B[2][1] * Ly[k]) + B[2][2] * Lz[k];
    }
}

// FaceExtraDirect increments (math/cfd/face_extra.h:135-149):
// du[k] = (gx[k]*dr.x + gy[k]*dr.y) + gz[k]*dr.z.
inline PState extrap(const PState& u, const PState& gx, const PState& gy, const PState& gz,
                     const Vec3& dr) {
    PState r;
    for (int k = 0; k < 6; ++k) {
        double du = (gx.arr()[k] * dr.x + gy.arr()[k] * dr.y) + gz.arr()[k] * dr.z;
        r.arr()[k] = u.arr()[k] + du;
    }
    return r;
}

// face.area(m_axial) and cell.volume(m_axial) (amr_faces.h:183-185,
// amr_cells.h:288-290).
inline double flux_area(const orc_smf& S, size_t s) { return S.axial ? S.m->area_alt[s] : S.m->area[s]; }
inline double flux_volume(const orc_smf& S, int c) { return S.axial ? S.m->vol_alt[c] : S.m->vol[c]; }

// Test-only planted mistakes of the axisymmetric step (0 none): 1 drops
// the source, 2 drops the predictor's 1/2 in it, 3 divides the update by
// the planar volume.
int g_plant_axial = 0;

// The axisymmetric pressure source: momentum.y += k * coeff * p * dt with
// coeff = volume / volume_alt (sm_fluid.cpp:166-169, 242-245, 347-350;
// k = 0.5 in the predictor).
inline void axial_source(const orc_smf& S, int c, double k, double p, QState& q) {
    if (!S.axial || g_plant_axial == 1) return;
    if (g_plant_axial == 2) k = 1.0;
    const double coeff = S.m->vol[c] / S.m->vol_alt[c];
    q.m.y += (k == 1.0 ? coeff : k * coeff) * p * S.dt;
}

inline double update_volume(const orc_smf& S, int c) {
    return g_plant_axial == 3 ? S.m->vol[c] : flux_volume(S, c);
}

// SmFluid::fluxes, first order (sm_fluid.cpp:127-174).
void fluxes_cell(orc_smf& S, int c) {
    const orc_mesh& m = *S.m;
    PState z_c = S.init[c];
    QState q_c = to_qstate(z_c);
    Flux flux{0, {0, 0, 0}, 0};
    for (int s8 = 0; s8 < 8; ++s8) {
        size_t s = 8 * (size_t)c + s8;
        int b = m.bnd[s];
        if (b == ORC_UNDEFINED) continue;
        Vec3 normal{m.fnx[s], m.fny[s], 0.0};
        PState z_n = !slot_is_boundary(b) ? S.init[m.adj[s]] : boundary_value(z_c, normal, b);
        PState zm = in_local(z_c, normal);
        PState zp = in_local(z_n, normal);
        Flux lf = num_flux(S.flux, zm, zp, S.eos);
        flux_to_global(lf, normal);
        for (int k = 0; k < 5; ++k) flux.arr()[k] = flux.arr()[k] + lf.arr()[k] * flux_area(S, s);
    }
    double coef = S.dt / update_volume(S, c);
    for (int k = 0; k < 5; ++k) q_c.arr()[k] = q_c.arr()[k] - coef * flux.arr()[k];
    axial_source(S, c, 1.0, z_c.p, q_c);
    S.next[c] = to_pstate(q_c, S.eos);
}

// SmFluid::fluxes_stage1, predictor (sm_fluid.cpp:176-253).
void stage1_cell(orc_smf& S, int c) {
    const orc_mesh& m = *S.m;
    if (S.wait[c] > 0) {
        S.half[c] = S.init[c];
        return;
    }
    Vec3 cell_c{m.cx[c], m.cy[c], 0.0};
    PState z_c = S.init[c];
    QState q_c = to_qstate(z_c);
    Flux flux{0, {0, 0, 0}, 0};
    for (int s8 = 0; s8 < 8; ++s8) {
        size_t s = 8 * (size_t)c + s8;
        int b = m.bnd[s];
        if (b == ORC_UNDEFINED) continue;
        Vec3 normal{m.fnx[s], m.fny[s], 0.0};
        Vec3 face_c{m.fcx[s], m.fcy[s], 0.0};
        // Only FaceExtraDirect::m is used here: it depends on the cell's own
        // gradient and dr_c = face_c - cell_c.
        PState zm = extrap(z_c, S.d_dx[c], S.d_dy[c], S.d_dz[c], face_c - cell_c);
        zm.e = S.eos.energy_rP(zm.rho, zm.p);
        if (is_bad(zm, S.eos)) zm = z_c;
        zm.v = rot_local(zm.v, normal);
        Flux lf = phys_flux(zm);
        flux_to_global(lf, normal);
        for (int k = 0; k < 5; ++k) flux.arr()[k] = flux.arr()[k] + lf.arr()[k] * flux_area(S, s);
    }
    double coef = 0.5 * S.dt / update_volume(S, c);
    for (int k = 0; k < 5; ++k) q_c.arr()[k] = q_c.arr()[k] - coef * flux.arr()[k];
    axial_source(S, c, 0.5, z_c.p, q_c);
    S.half[c] = to_pstate(q_c, S.eos);
    if (is_bad(S.half[c], S.eos)) S.half[c] = z_c;
}

// SmFluid::fluxes_stage2, corrector (sm_fluid.cpp:255-366).  `wait` is read
// from the pre-pass array and written to wait_next (the reference reads a
// neighbour's wait while other threads update theirs, sm_fluid.cpp:293 vs
// :357/:362; the double buffer fixes the pre-pass semantics).
void stage2_cell(orc_smf& S, int c) {
    const orc_mesh& m = *S.m;
    Vec3 cell_c{m.cx[c], m.cy[c], 0.0};
    PState z_c = S.init[c];
    PState z_ch = S.half[c];
    QState q_c = to_qstate(z_c);
    Flux flux{0, {0, 0, 0}, 0};
    for (int s8 = 0; s8 < 8; ++s8) {
        size_t s = 8 * (size_t)c + s8;
        int b = m.bnd[s];
        if (b == ORC_UNDEFINED) continue;
        Vec3 normal{m.fnx[s], m.fny[s], 0.0};
        Vec3 face_c{m.fcx[s], m.fcy[s], 0.0};
        int nb = m.adj[s];   // the cell itself on boundary slots
        PState z_n, z_nh;
        bool bound = slot_is_boundary(b);
        if (!bound) {
            z_n = S.init[nb];
            z_nh = S.half[nb];
        } else {
            z_n = boundary_value(z_c, normal, b);
            z_nh = boundary_value(z_ch, normal, b);
        }
        (void)z_n;
        PState zm, zp;
        bool robust = S.wait[c] > 0 || S.wait[nb] > 0;
        if (robust) {
            zm = z_ch;
            zp = z_nh;
        } else {
            Vec3 neib_c = (b == ORC_INNER) ? Vec3{m.cx[nb], m.cy[nb], 0.0} : symm_point(m, (int)s, cell_c);
            zm = extrap(z_ch, S.d_dx[c], S.d_dy[c], S.d_dz[c], face_c - cell_c);
            zm.e = S.eos.energy_rP(zm.rho, zm.p);
            if (is_bad(zm, S.eos)) zm = z_ch;
            if (!bound) {
                zp = extrap(z_nh, S.d_dx[nb], S.d_dy[nb], S.d_dz[nb], face_c - neib_c);
                zp.e = S.eos.energy_rP(zp.rho, zp.p);
                if (is_bad(zp, S.eos)) zp = z_nh;
            } else {
                zp = boundary_value(zm, normal, b);
            }
        }
        zm.v = rot_local(zm.v, normal);
        zp.v = rot_local(zp.v, normal);
        Flux lf = robust ? hll(zm, zp, S.eos) : num_flux(S.flux, zm, zp, S.eos);
        flux_to_global(lf, normal);
        for (int k = 0; k < 5; ++k) flux.arr()[k] = flux.arr()[k] + lf.arr()[k] * flux_area(S, s);
    }
    double coef = S.dt / update_volume(S, c);
    for (int k = 0; k < 5; ++k) q_c.arr()[k] = q_c.arr()[k] - coef * flux.arr()[k];
    axial_source(S, c, 1.0, z_ch.p, q_c);
    PState zn = to_pstate(q_c, S.eos);
    if (is_bad(zn, S.eos)) {
        S.next[c] = z_c;
        S.wait_next[c] = 3;
    } else {
        S.next[c] = zn;
        S.wait_next[c] = S.wait[c] > 0 ? S.wait[c] - 1 : S.wait[c];
    }
}

// SmFluid::compute_dt (sm_fluid.cpp:104-114) with threads::min (threads.h:292-352).
double compute_dt(orc_smf& S, int nthreads) {
    const orc_mesh& m = *S.m;
    double res = std::numeric_limits<double>::max();
    #pragma omp parallel for num_threads(nthreads) schedule(static) reduction(min:res)
    for (int c = 0; c < m.n; ++c) {
        const PState& z = S.init[c];
        double cs = S.eos.sound_speed_rP(z.rho, z.p);
        double t = incircle(m, c) / (norm(z.v) + cs);
        if (t < res) res = t;
    }
    return smin(S.cfl * res, S.max_dt);
}

}  // namespace

extern "C" {

orc_smf* orc_smf_create(orc_mesh* m, double gamma, double P0, double e0,
                        int flux, int limiter, int acc, double cfl) {
    orc_smf* S = new orc_smf();
    S->m = m;
    S->eos = {gamma, P0, e0};
    S->flux = flux;
    S->limiter = limiter;
    S->acc = std::min(std::max(1, acc), 2);             // set_accuracy, sm_fluid.cpp:41-43
    S->cfl = std::max(0.0, std::min(cfl, 1.0));         // set_CFL, sm_fluid.cpp:37-39
    S->dt = NAN;
    S->max_dt = std::numeric_limits<double>::max();     // sm_fluid.cpp:22
    int n = m->n;
    PState zz; zero(zz);
    S->init.assign(n, zz); S->half.assign(n, zz); S->next.assign(n, zz);
    S->d_dx.assign(n, zz); S->d_dy.assign(n, zz); S->d_dz.assign(n, zz);
    S->wait.assign(n, 0); S->wait_next.assign(n, 0);
    return S;
}

void orc_smf_free(orc_smf* s) { delete s; }
void orc_smf_set_max_dt(orc_smf* s, double max_dt) { s->max_dt = max_dt; }

void orc_smf_plant_axial(int k) { g_plant_axial = k; }

int orc_smf_set_axial(orc_smf* s, int axial) {
    if (axial && !s->m->axial) return -1;   // no alternative geometry on this mesh
    s->axial = axial != 0;
    return 0;
}

void orc_smf_set_state(orc_smf* s, const double* rho, const double* u, const double* v,
                       const double* p, const double* e) {
    for (int c = 0; c < s->m->n; ++c) s->init[c] = PState{rho[c], {u[c], v[c], 0.0}, p[c], e[c]};
}

void orc_smf_get_state(const orc_smf* s, double* rho, double* u, double* v, double* p, double* e) {
    for (int c = 0; c < s->m->n; ++c) {
        const PState& z = s->init[c];
        rho[c] = z.rho; u[c] = z.v.x; v[c] = z.v.y; p[c] = z.p; e[c] = z.e;
    }
}

void orc_smf_get_wait(const orc_smf* s, int* wait) {
    std::memcpy(wait, s->wait.data(), sizeof(int) * s->m->n);
}

void orc_smf_set_wait(orc_smf* s, const int* wait) {
    std::memcpy(s->wait.data(), wait, sizeof(int) * s->m->n);
}

void orc_smf_get_grad(const orc_smf* s, double* dx, double* dy) {
    int n = s->m->n;
    for (int c = 0; c < n; ++c) {
        const PState& gx = s->d_dx[c];
        const PState& gy = s->d_dy[c];
        dx[c] = gx.rho; dx[n + c] = gx.v.x; dx[2 * n + c] = gx.v.y; dx[3 * n + c] = gx.p;
        dy[c] = gy.rho; dy[n + c] = gy.v.x; dy[2 * n + c] = gy.v.y; dy[3 * n + c] = gy.p;
    }
}

double orc_smf_compute_dt(orc_smf* s, int nthreads) { return compute_dt(*s, nthreads); }

void orc_smf_compute_grad(orc_smf* s, int nthreads) {
    int n = s->m->n;
    #pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int c = 0; c < n; ++c) {
        Grad g;
        grad_cell(*s, c, g);
        s->d_dx[c] = g.x; s->d_dy[c] = g.y; s->d_dz[c] = g.z;
    }
}

// SmFluid::update (sm_fluid.cpp:81-102).
int orc_smf_update(orc_smf* s, int nthreads) {
    int n = s->m->n;
    int err = 0;
    s->dt = compute_dt(*s, nthreads);
    if (s->acc == 1) {
        #pragma omp parallel for num_threads(nthreads) schedule(static) reduction(|:err)
        for (int c = 0; c < n; ++c) {
            g_numeric_error = 0;
            fluxes_cell(*s, c);
            err |= g_numeric_error;
        }
    } else {
        orc_smf_compute_grad(s, nthreads);
        #pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int c = 0; c < n; ++c) stage1_cell(*s, c);
        #pragma omp parallel for num_threads(nthreads) schedule(static) reduction(|:err)
        for (int c = 0; c < n; ++c) {
            g_numeric_error = 0;
            stage2_cell(*s, c);
            err |= g_numeric_error;
        }
        std::swap(s->wait, s->wait_next);
    }
    std::swap(s->init, s->next);   // SmFluid::swap (sm_fluid.cpp:368-370)
    return err;
}

double orc_smf_dt(const orc_smf* s) { return s->dt; }

void orc_smf_totals(const orc_smf* s, double* out) {
    double t[4] = {0, 0, 0, 0};
    for (int c = 0; c < s->m->n; ++c) {
        QState q = to_qstate(s->init[c]);
        double V = flux_volume(*s, c);
        t[0] += V * q.rho; t[1] += V * q.m.x; t[2] += V * q.m.y; t[3] += V * q.E;
    }
    for (int k = 0; k < 4; ++k) out[k] = t[k];
}

// ---- point functions ------------------------------------------------------
int orc_flux(int kind, const double* zL, const double* zR, const double* eos, double* F) {
    PState L{zL[0], {zL[1], zL[2], zL[3]}, zL[4], zL[5]};
    PState R{zR[0], {zR[1], zR[2], zR[3]}, zR[4], zR[5]};
    Eos e{eos[0], eos[1], eos[2]};
    g_numeric_error = 0;
    Flux f = num_flux(kind, L, R, e);
    for (int k = 0; k < 5; ++k) F[k] = f.arr()[k];
    return g_numeric_error;
}

double orc_limiter2(int kind, double a, double b) { return lim_binary(kind, a, b); }
double orc_limiter1(int kind, double r) { return lim_unary(kind, r); }

void orc_riemann(double rL, double uL, double pL, double gL, double p0L,
                 double rR, double uR, double pR, double gR, double p0R, double* out) {
    RSol s = riemann_solve(rL, uL, pL, gL, p0L, rR, uR, pR, gR, p0R);
    out[0] = s.rho; out[1] = s.Uf; out[2] = s.Pf; out[3] = s.P; out[4] = s.U; out[5] = s.conv;
}

// RiemannSolver::compute + density/velocity/pressure (riemann.cpp:494-613).
void orc_riemann_sample(double rL, double uL, double pL, double rR, double uR, double pR,
                        double g, double p0, double xi, double* out) {
    double cL = eos_sound(rL, pL, g, p0), cR = eos_sound(rR, pR, g, p0);
    SolPU UP = contact_p(rL, uL, pL, cL, g, p0, rR, uR, pR, cR, g, p0);
    double P = UP.P, U = UP.U, DL1, DL2, DR1, DR2, rl, rr;
    if (pL < P) {
        DL1 = DL2 = shock_wave_speed(rL, uL, pL, U, P);
        rl = shock_wave_density(rL, uL, pL, U, P);
    } else {
        double cl = rarefaction_sound_L(U, uL, cL, g);
        DL1 = uL - cL; DL2 = U - cl; rl = eos_density(P, cl, g, p0);
    }
    if (pR < P) {
        DR1 = DR2 = shock_wave_speed(rR, uR, pR, U, P);
        rr = shock_wave_density(rR, uR, pR, U, P);
    } else {
        double cr = rarefaction_sound_R(U, uR, cR, g);
        DR1 = U + cr; DR2 = uR + cR; rr = eos_density(P, cr, g, p0);
    }
    double rho, u, p;
    if (xi < DL1) { rho = rL; }
    else if (xi < DL2) { rho = density_lfan(rL, uL, cL, g, xi); }
    else if (xi < U) { rho = rl; }
    else if (xi < DR1) { rho = rr; }
    else if (xi < DR2) { rho = density_rfan(rR, uR, cR, g, xi); }
    else { rho = rR; }
    if (xi < DL1) { u = uL; p = pL; }
    else if (xi < DL2) { u = velocity_lfan(uL, cL, g, xi); p = pressure_lfan(uL, pL, cL, g, p0, xi); }
    else if (xi < DR1) { u = U; p = P; }
    else if (xi < DR2) { u = velocity_rfan(uR, cR, g, xi); p = pressure_rfan(uR, pR, cR, g, p0, xi); }
    else { u = uR; p = pR; }
    out[0] = rho; out[1] = u; out[2] = p;
}

void orc_eos(const double* eos, double rho, double e, double p, double* out) {
    Eos E{eos[0], eos[1], eos[2]};
    out[0] = E.pressure_re(rho, e);
    out[1] = E.energy_rP(rho, p);
    out[2] = E.sound_speed_rP(rho, p);
    out[3] = E.min_pressure();
}

// ---- synthetic inputs -----------------------------------------------------
double orc_splitmix_uniform(uint64_t* state) {
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * 0x1.0p-53;
}

void orc_init_quadrants(const orc_mesh* m, uint64_t seed, double gamma,
                        double* rho, double* u, double* v, double* p, double* e) {
    uint64_t st = seed;
    double R[4], U[4], V[4], P[4];
    for (int q = 0; q < 4; ++q) {   // quadrants 1..4, draw order rho, u, v, p
        R[q] = 0.1 + 1.9 * orc_splitmix_uniform(&st);
        U[q] = -1.0 + 2.0 * orc_splitmix_uniform(&st);
        V[q] = -1.0 + 2.0 * orc_splitmix_uniform(&st);
        P[q] = 0.1 + 1.9 * orc_splitmix_uniform(&st);
    }
    Eos E{gamma, 0.0, 0.0};
    for (int c = 0; c < m->n; ++c) {
        double x = m->cx[c], y = m->cy[c];
        int q;   // Riemann2D::density (phys/tests/test_2D.cpp:72-83)
        if (x >= 0.5 && y >= 0.5) q = 0;
        else if (x <= 0.5 && y >= 0.5) q = 1;
        else if (x <= 0.5 && y <= 0.5) q = 2;
        else q = 3;
        rho[c] = R[q]; u[c] = U[q]; v[c] = V[q]; p[c] = P[q];
        e[c] = E.energy_rP(R[q], P[q]);
    }
}

}  // extern "C"

#include "oracle_amr.hpp"
#include "oracle_mm.hpp"
#include "oracle_transfer.hpp"
#include "oracle_3d.hpp"
