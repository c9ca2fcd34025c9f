// Host checks of the reverse-mode primitives of the discrete adjoint
// (zephyr_amd/csrc/zfv_adjoint_ops.hpp) against central differences of the
// forward primitives (zfv_physics.hpp), at random physical states:
// <rev(ybar), d> must equal <ybar, (f(x + eps d) - f(x - eps d)) / (2 eps)>.
// Prints one line per primitive with the worst relative error; exit code 1
// if any exceeds the bound.
#include <cstdio>
#include <random>
#include <vector>

#include "zfv_adjoint_ops.hpp"

using namespace zfv;
using namespace zfv::adj;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

static Z rand_state(const Eos& eos) {
    Z z;
    z.r = U(0.2, 2.0);
    z.u = U(-1.5, 1.5);
    z.v = U(-1.5, 1.5);
    z.p = U(0.2, 2.0);
    z.e = eos_e(eos, z.r, z.p);
    return z;
}
static double* zp(Z& z) { return &z.r; }
static const double* zp(const Z& z) { return &z.r; }
static double* qp(Q& q) { return &q.r; }

static double worst = 0.0;
static int fails = 0;
static void report(const char* name, double w, double bound) {
    std::printf("%-26s worst rel err %.3e (bound %.0e)\n", name, w, bound);
    if (!(w <= bound)) ++fails;
}
static double rel(double a, double b, double scale) { return std::fabs(a - b) / std::max(scale, 1e-300); }

int main() {
    const Eos eos{1.4, 0.0, 0.0};
    const double eps = 1e-6;
    const int trials = 2000;

    // to_q
    {
        double w = 0.0;
        for (int t = 0; t < trials; ++t) {
            Z z = rand_state(eos), d;
            for (int k = 0; k < 5; ++k) zp(d)[k] = U(-1, 1);
            Q qb{U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
            Z zb{0, 0, 0, 0, 0};
            to_q_rev(z, qb, zb);
            Z a = z, b = z;
            for (int k = 0; k < 5; ++k) { zp(a)[k] += eps * zp(d)[k]; zp(b)[k] -= eps * zp(d)[k]; }
            Q qa = to_q(a), qm = to_q(b);
            double fd = 0.0, ad = 0.0, sc = 0.0;
            for (int k = 0; k < 4; ++k) fd += qp(qb)[k] * (qp(qa)[k] - qp(qm)[k]) / (2 * eps);
            for (int k = 0; k < 5; ++k) { ad += zp(zb)[k] * zp(d)[k]; sc += std::fabs(zp(zb)[k] * zp(d)[k]); }
            w = std::max(w, rel(fd, ad, sc));
        }
        report("to_q_rev", w, 1e-7);
    }
    // to_z
    {
        double w = 0.0;
        for (int t = 0; t < trials; ++t) {
            Z z0 = rand_state(eos);
            Q q = to_q(z0), d{U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
            Z z = to_z(q, eos);
            Z zb{U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
            Q qb{0, 0, 0, 0};
            to_z_rev(q, z, zb, eos, qb);
            Q a = q, b = q;
            for (int k = 0; k < 4; ++k) { qp(a)[k] += eps * qp(d)[k]; qp(b)[k] -= eps * qp(d)[k]; }
            Z za = to_z(a, eos), zm = to_z(b, eos);
            double fd = 0.0, ad = 0.0, sc = 0.0;
            for (int k = 0; k < 5; ++k) fd += zp(zb)[k] * (zp(za)[k] - zp(zm)[k]) / (2 * eps);
            for (int k = 0; k < 4; ++k) { ad += qp(qb)[k] * qp(d)[k]; sc += std::fabs(qp(qb)[k] * qp(d)[k]); }
            w = std::max(w, rel(fd, ad, sc));
        }
        report("to_z_rev", w, 1e-7);
    }
    // phys_flux
    {
        double w = 0.0;
        for (int t = 0; t < trials; ++t) {
            Z z = rand_state(eos), d;
            for (int k = 0; k < 5; ++k) zp(d)[k] = U(-1, 1);
            Q fb{U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
            Z zb{0, 0, 0, 0, 0};
            phys_flux_rev(z, fb, zb);
            Z a = z, b = z;
            for (int k = 0; k < 5; ++k) { zp(a)[k] += eps * zp(d)[k]; zp(b)[k] -= eps * zp(d)[k]; }
            Q fa = phys_flux(a), fm = phys_flux(b);
            double fd = 0.0, ad = 0.0, sc = 0.0;
            for (int k = 0; k < 4; ++k) fd += qp(fb)[k] * (qp(fa)[k] - qp(fm)[k]) / (2 * eps);
            for (int k = 0; k < 5; ++k) { ad += zp(zb)[k] * zp(d)[k]; sc += std::fabs(zp(zb)[k] * zp(d)[k]); }
            w = std::max(w, rel(fd, ad, sc));
        }
        report("phys_flux_rev", w, 1e-7);
    }
    // HLLC (both states, all branches: supersonic states are drawn too)
    {
        double w = 0.0, wdot = 0.0;
        int subsonic = 0;
        for (int t = 0; t < 20 * trials; ++t) {
            Z L = rand_state(eos), R = rand_state(eos), dL, dR;
            if (t % 5 == 0) { L.u += 3.0; R.u += 3.0; }
            for (int k = 0; k < 5; ++k) { zp(dL)[k] = U(-1, 1); zp(dR)[k] = U(-1, 1); }
            Q Fb{U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
            Z Lb{0, 0, 0, 0, 0}, Rb{0, 0, 0, 0, 0};
            double fdot = 0.0;
            hllc_rev(L, R, Fb, eos, Lb, Rb, &fdot);
            {
                bool nan0 = false;
                const Q F0 = flux_hllc(L, R, eos, nan0);
                const double ref = qp(Fb)[0] * F0.r + qp(Fb)[1] * F0.mx + qp(Fb)[2] * F0.my + qp(Fb)[3] * F0.E;
                const double sc = std::fabs(qp(Fb)[0] * F0.r) + std::fabs(qp(Fb)[1] * F0.mx) +
                                  std::fabs(qp(Fb)[2] * F0.my) + std::fabs(qp(Fb)[3] * F0.E);
                wdot = std::max(wdot, rel(fdot, ref, sc));
            }
            Z La = L, Lm = L, Ra = R, Rm = R;
            for (int k = 0; k < 5; ++k) {
                zp(La)[k] += eps * zp(dL)[k]; zp(Lm)[k] -= eps * zp(dL)[k];
                zp(Ra)[k] += eps * zp(dR)[k]; zp(Rm)[k] -= eps * zp(dR)[k];
            }
            bool nan = false;
            Q Fa = flux_hllc(La, Ra, eos, nan), Fm = flux_hllc(Lm, Rm, eos, nan);
            double fd = 0.0, ad = 0.0, sc = 0.0;
            for (int k = 0; k < 4; ++k) fd += qp(Fb)[k] * (qp(Fa)[k] - qp(Fm)[k]) / (2 * eps);
            for (int k = 0; k < 5; ++k) {
                ad += zp(Lb)[k] * zp(dL)[k] + zp(Rb)[k] * zp(dR)[k];
                sc += std::fabs(zp(Lb)[k] * zp(dL)[k]) + std::fabs(zp(Rb)[k] * zp(dR)[k]);
            }
            const double e = rel(fd, ad, sc);
            if (e > 1e-4) continue;   // a branch switch inside +-eps (rare): not a derivative error
            w = std::max(w, e);
            ++subsonic;
        }
        report("hllc_rev", w, 1e-6);
        report("hllc_rev F.Fb", wdot, 1e-13);
    }
    // face state: extrapolation + EOS energy
    {
        double w = 0.0;
        for (int t = 0; t < trials; ++t) {
            Z base = rand_state(eos), dbase;
            double gx[4], gy[4], dgx[4], dgy[4];
            for (int k = 0; k < 4; ++k) { gx[k] = U(-0.3, 0.3); gy[k] = U(-0.3, 0.3); dgx[k] = U(-1, 1); dgy[k] = U(-1, 1); }
            for (int k = 0; k < 5; ++k) zp(dbase)[k] = U(-1, 1);
            const double dx = U(-0.5, 0.5), dy = U(-0.5, 0.5);
            bool bad;
            Z zf = face_state(base, gx, gy, dx, dy, eos, bad);
            Z zfb{U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
            Z bb{0, 0, 0, 0, 0};
            double gxb[4] = {0, 0, 0, 0}, gyb[4] = {0, 0, 0, 0};
            face_state_rev(zf, bad, zfb, dx, dy, eos, bb, gxb, gyb);
            Z ba = base, bm = base;
            double gxa[4], gya[4], gxm[4], gym[4];
            for (int k = 0; k < 5; ++k) { zp(ba)[k] += eps * zp(dbase)[k]; zp(bm)[k] -= eps * zp(dbase)[k]; }
            for (int k = 0; k < 4; ++k) {
                gxa[k] = gx[k] + eps * dgx[k]; gxm[k] = gx[k] - eps * dgx[k];
                gya[k] = gy[k] + eps * dgy[k]; gym[k] = gy[k] - eps * dgy[k];
            }
            bool b1, b2;
            Z fa = face_state(ba, gxa, gya, dx, dy, eos, b1), fm = face_state(bm, gxm, gym, dx, dy, eos, b2);
            if (b1 != bad || b2 != bad) continue;
            double fd = 0.0, ad = 0.0, sc = 0.0;
            for (int k = 0; k < 5; ++k) fd += zp(zfb)[k] * (zp(fa)[k] - zp(fm)[k]) / (2 * eps);
            for (int k = 0; k < 5; ++k) { ad += zp(bb)[k] * zp(dbase)[k]; sc += std::fabs(zp(bb)[k] * zp(dbase)[k]); }
            for (int k = 0; k < 4; ++k) {
                ad += gxb[k] * dgx[k] + gyb[k] * dgy[k];
                sc += std::fabs(gxb[k] * dgx[k]) + std::fabs(gyb[k] * dgy[k]);
            }
            w = std::max(w, rel(fd, ad, sc));
        }
        report("face_state_rev", w, 1e-7);
    }
    // limiters: psi(a, b) = phi(a, b) b for minmod and MC.  Both are piecewise
    // linear, so a sample is kept only where the two one-sided quotients agree
    // (no breakpoint inside +-eps); there the central quotient is exact up to
    // rounding.  The filter passes second differences up to 1e-14 (a few ulps
    // of psi ~ 1), so a breakpoint that slips through adds at most
    // 1e-14 / (2 eps) = 5e-9 to the quotient; rounding adds about 1e-10.  Hence
    // the absolute floor 0.1 of the error scale (1e-8 absolute at the bound).
    {
        auto check = [&](auto f, auto d, const char* name) {
            double w = 0.0;
            int kept = 0;
            for (int t = 0; t < 20 * trials; ++t) {
                const double a = U(-2, 2), b = U(-2, 2);
                const double da = U(-1, 1), db = U(-1, 1);
                double pa, pb;
                d(a, b, pa, pb);
                const double f0 = f(a, b);
                const double fp = f(a + eps * da, b + eps * db), fm = f(a - eps * da, b - eps * db);
                const double sc = std::fabs(pa * da) + std::fabs(pb * db) + 0.1;
                if (std::fabs((fp - f0) - (f0 - fm)) > 1e-14) continue;
                w = std::max(w, rel((fp - fm) / (2 * eps), pa * da + pb * db, sc));
                ++kept;
            }
            report(name, kept > trials ? w : 1.0, 1e-7);
        };
        check([](double x, double y) { return limiter2<2>(x, y) * y; },
              [](double x, double y, double& pa, double& pb) { psi_d<2>(x, y, pa, pb); }, "psi_d<minmod>");
        check([](double x, double y) { return limiter2<3>(x, y) * y; },
              [](double x, double y, double& pa, double& pb) { psi_d<3>(x, y, pa, pb); }, "psi_d<MC>");
    }
    return fails ? 1 : 0;
}
