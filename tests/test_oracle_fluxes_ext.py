"""Known answers for the oracle's HLLC_M, HLLC_LM, CIR1 and CIR2 fluxes
(math/cfd/flux/hllc.cpp:427-572, cir.cpp:16-216), fixed by the mathematics
rather than by the formulas themselves:

* consistency F(z, z) = Flux(z), and the supersonic early-outs;
* mirror antisymmetry (x -> -x swaps the sides and negates the normal flux);
* HLLC_M leaves HLLC's mass and normal-momentum fluxes bit for bit (only the
  tangential velocity and energy of the star states change) and equals HLLC
  to round-off when the tangential velocities agree;
* HLLC_LM is HLLC bit for bit above Ma = 0.1, and for a colliding
  symmetric pair of states (S_C = 0, S_L = -S_R) its normal-momentum flux
  beyond (F_L + F_R) / 2 is phi = sin(pi/2 Ma / 0.1) times HLLC's, with phi
  from numpy;
* CIR1 and CIR2: with a supersonic face state every eigenvalue is positive,
  |A| = A, and the flux is (F_L + F_R)/2 + A (q_L - q_R)/2 with A the flux
  Jacobian at the arithmetic face state, here by central differences of the
  physical flux (in (rho, m, E), and in (rho, m, E, p) with p independent
  for CIR2); for a subsonic face state |A| = V |Lambda| V^-1 from numpy's
  eigen-decomposition of that Jacobian; for small jumps CIR2 tends to CIR1.
A wrong entry of either eigenvector matrix, a swapped eigenvalue or a wrong
phi fails one of these.  The reference forms the CIR products with Eigen;
the restatement's summation order is parity unpinned, which these
tolerances (1e-6 and looser) do not distinguish.
"""
import numpy as np
import pytest

GAMMA = 1.4
HLLC, HLLC_M, HLLC_LM, CIR1, CIR2 = 3, 4, 5, 6, 7
NEW = {"HLLC_M": HLLC_M, "HLLC_LM": HLLC_LM, "CIR1": CIR1, "CIR2": CIR2}


def z6(rho, u, v, p, gamma=GAMMA, P0=0.0, w=0.0):
    return np.array([rho, u, v, w, p, (p + gamma * P0) / ((gamma - 1.0) * rho)])


def phys(z, gamma=GAMMA, P0=0.0):
    rho, u, v, w, p, e = z
    E = rho * (e + 0.5 * (u * u + v * v + w * w))
    return np.array([rho * u, rho * u * u + p, rho * u * v, rho * u * w, (E + p) * u])


def cons(z):
    rho, u, v, w, p, e = z
    return np.array([rho, rho * u, rho * v, rho * w, rho * (e + 0.5 * (u * u + v * v + w * w))])


def flux_of_q(q, gamma=GAMMA, P0=0.0):
    rho, mx, my, mz, E = q
    u, v, w = mx / rho, my / rho, mz / rho
    e = E / rho - 0.5 * (u * u + v * v + w * w)
    p = (gamma - 1.0) * rho * e - gamma * P0
    return phys(np.array([rho, u, v, w, p, e]), gamma, P0)


def flux_of_qp(qp):
    """The Euler flux with the pressure as an independent sixth variable."""
    rho, mx, my, mz, E, p = qp
    u = mx / rho
    return np.array([mx, mx * u + p, my * u, mz * u, (E + p) * u])


def fd_jacobian(f, x, rel=1e-6):
    J = np.zeros((5, len(x)))
    for k in range(len(x)):
        h = rel * max(abs(x[k]), 1.0)
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        J[:, k] = (f(xp) - f(xm)) / (2 * h)
    return J


def face_state(zL, zR, gamma=GAMMA, P0=0.0):
    zE = 0.5 * (zL + zR)
    zE[5] = (zE[4] + gamma * P0) / ((gamma - 1.0) * zE[0])
    return zE


STATES = [(z6(1.0, 0.3, -0.2, 1.0), z6(0.7, 0.1, 0.4, 0.6)),
          (z6(0.2, -0.8, 0.5, 0.3), z6(1.4, 0.2, -0.1, 1.7)),
          (z6(2.0, 0.0, 0.0, 1.0), z6(1.0, 0.0, 0.0, 1.0))]


@pytest.mark.parametrize("name", sorted(NEW))
def test_consistency_and_supersonic(oracle, name):
    kind = NEW[name]
    for z in (z6(1.0, 0.3, -0.2, 1.0), z6(0.3, -2.5, 1.0, 0.4), z6(5.0, 0.0, 0.1, 7.0)):
        F, err = oracle.flux(kind, z, z)
        assert err == 0
        np.testing.assert_allclose(F, phys(z), rtol=2e-15, atol=2e-15)
        if name.startswith("CIR"):
            assert np.array_equal(F, phys(z)), "dq = 0 exactly"
    if name.startswith("HLLC"):
        zL, zR = z6(1.0, 5.0, 0.3, 1.0), z6(0.5, 4.0, -0.2, 0.8)   # S_L > 0
        F, _ = oracle.flux(kind, zL, zR)
        assert np.array_equal(F, phys(zL))
        F, _ = oracle.flux(kind, z6(1.0, -5.0, 0.1, 1.0), z6(0.5, -4.0, 0.2, 0.8))   # S_R < 0
        assert np.array_equal(F, phys(z6(0.5, -4.0, 0.2, 0.8)))


@pytest.mark.parametrize("name", sorted(NEW))
def test_mirror_antisymmetry(oracle, name):
    kind = NEW[name]
    for zL, zR in STATES:
        F, _ = oracle.flux(kind, zL, zR)
        mL, mR = zL.copy(), zR.copy()
        mL[1], mR[1] = -zL[1], -zR[1]
        G, _ = oracle.flux(kind, mR, mL)
        expect = -F
        expect[2], expect[3] = F[2] * -1.0, F[3] * -1.0
        expect[1] = F[1]
        np.testing.assert_allclose(G, expect, rtol=1e-13, atol=1e-13 * np.abs(F).max())


def test_hllc_m_against_hllc(oracle):
    for zL, zR in STATES[:2]:
        F, _ = oracle.flux(HLLC, zL, zR)
        M, _ = oracle.flux(HLLC_M, zL, zR)
        assert F[0] == M[0] and F[1] == M[1], "mass and normal momentum must be HLLC's bit for bit"
        assert abs(F[2] - M[2]) > 1e-6, "different tangential velocities average in HLLC_M"
        same_t = zR.copy()
        same_t[2] = zL[2]
        F, _ = oracle.flux(HLLC, zL, same_t)
        M, _ = oracle.flux(HLLC_M, zL, same_t)
        np.testing.assert_allclose(M, F, rtol=1e-13, atol=1e-14)


def test_hllc_lm_mach_scaling(oracle):
    for zL, zR in STATES[:2]:   # Ma > 0.1: phi = 1
        F, _ = oracle.flux(HLLC, zL, zR)
        L, _ = oracle.flux(HLLC_LM, zL, zR)
        assert np.array_equal(F, L)
    for frac in (0.02, 0.05, 0.08):
        # colliding symmetric states: S_C = 0 and S_L = -S_R, so the normal
        # momentum flux is (F_L + F_R)/2 = rho u^2 + p plus the outer-wave
        # term, which HLLC_LM scales by phi
        c = np.sqrt(GAMMA * 1.0 / 1.0)
        u = frac * c
        zL, zR = z6(1.0, u, 0.0, 1.0), z6(1.0, -u, 0.0, 1.0)
        F, _ = oracle.flux(HLLC, zL, zR)
        L, _ = oracle.flux(HLLC_LM, zL, zR)
        phi = np.sin(0.5 * np.pi * frac / 0.1)
        m0 = 1.0 * u * u + 1.0
        assert abs(F[1] - m0) > 1e-3
        assert abs((L[1] - m0) - phi * (F[1] - m0)) <= 1e-12 * abs(F[1] - m0), (frac, L[1], F[1], phi)


def cir_expected(zL, zR, extended):
    zE = face_state(zL, zR)
    qE = cons(zE)
    if extended:
        x = np.concatenate([qE, [zE[4]]])
        A = fd_jacobian(flux_of_qp, x)
        dq = np.concatenate([cons(zL) - cons(zR), [zL[4] - zR[4]]])
    else:
        A = fd_jacobian(flux_of_q, qE)
        dq = cons(zL) - cons(zR)
    return A, dq


@pytest.mark.parametrize("kind,extended", [(CIR1, False), (CIR2, True)])
def test_cir_supersonic_is_the_jacobian(oracle, kind, extended):
    for zL, zR in ((z6(1.0, 2.2, 0.3, 1.0), z6(0.8, 2.0, -0.1, 0.7)),
                   (z6(0.5, -3.0, 0.2, 0.4), z6(0.6, -2.6, 0.5, 0.5))):
        A, dq = cir_expected(zL, zR, extended)
        zE = face_state(zL, zR)
        sign = 1.0 if zE[1] > 0 else -1.0   # every eigenvalue of one sign: |A| = +-A
        expect = 0.5 * (phys(zL) + phys(zR)) + 0.5 * sign * (A @ dq)
        F, err = oracle.flux(kind, zL, zR)
        assert err == 0
        np.testing.assert_allclose(F, expect, rtol=1e-6, atol=1e-6 * np.abs(expect).max())


def test_cir1_subsonic_is_abs_jacobian(oracle):
    for zL, zR in STATES[:2] + [(z6(1.0, 0.2, 0.1, 1.0), z6(0.9, 0.25, 0.05, 0.95))]:
        A, dq = cir_expected(zL, zR, False)
        lam, V = np.linalg.eig(A)
        absA = np.real(V @ np.diag(np.abs(np.real(lam))) @ np.linalg.inv(V))
        expect = 0.5 * (phys(zL) + phys(zR)) + 0.5 * absA @ dq
        F, _ = oracle.flux(CIR1, zL, zR)
        np.testing.assert_allclose(F, expect, rtol=1e-6, atol=1e-6 * np.abs(expect).max())


def test_cir2_tends_to_cir1(oracle):
    base = z6(1.0, 0.3, -0.2, 1.0)
    for eps in (1e-3, 1e-4):
        zR = z6(1.0 + eps, 0.3 - eps, -0.2 + 0.5 * eps, 1.0 + 2 * eps)
        F1, _ = oracle.flux(CIR1, base, zR)
        F2, _ = oracle.flux(CIR2, base, zR)
        diff = np.abs(F1 - F2).max()
        jump = np.abs(F1 - phys(base)).max()
        assert diff <= 10 * eps * jump + 1e-14, (eps, diff, jump)


def test_new_fluxes_step_conservative(oracle):
    """A periodic box keeps mass, momentum and energy with every new flux,
    at both orders."""
    for kind in NEW.values():
        for acc in (1, 2):
            m = oracle.Mesh.rectangle(16, 12, bounds=(3, 3, 3, 3))
            s = oracle.SmFluid(m, acc=acc, flux=kind, limiter="MC", cfl=0.4)
            s.set_state(*m.quadrant_state(9))
            t0 = s.totals()
            for _ in range(5):
                assert s.update(2) == 0
            np.testing.assert_allclose(s.totals(), t0, rtol=1e-13, atol=1e-13)
