"""Pins of the oracle's composite-Riemann-flux chain (CrpFlux classic /
inverse, the HLLC wave configurations they are built on, PState::split and
calc_crp_flux's dispatch) by answers the mathematics fixes, independent of
the restatement's code:

* a contact translated at uniform velocity u and pressure P is an exact
  solution: the time-averaged flux through the face over dt is the physical
  flux of the state that crosses it, or, once the interface at distance
  delta reaches the face after t* = delta / |u|, f_far - (q_far - q_near)
  delta / dt (the Rankine-Hugoniot relation f - u q continuous across the
  contact).  Each CRP sub-problem must return it to round-off wherever its
  wave pattern is resolved (the CFL-limited steps of the scheme; in the
  clear-to-mix branch the reference's construction is exact only when the
  face itself carries no density jump, so those samples take equal states
  on both sides of the face);
* the composite flux is mirror-antisymmetric: swapping the two cells and
  reversing x negates the flux (the classic problem of one orientation is
  the inverse problem of the other);
* PState::split keeps mass and internal energy: alpha_A rho_A + (1 -
  alpha_A) rho_B = rho and the mass-weighted energies add up to e, and the
  mixture part closes back to the same P and T;
* a two-material interface at rest in pressure equilibrium stays at rest
  with composite fluxes on, and no material crosses it.

A mistake planted in the oracle (a swapped classic / inverse call, a flipped
wave-pattern branch, a wrong star state, a wrong split energy) must make at
least one of these checks fail (test_planted_mistakes_are_caught).
"""
import numpy as np
import pytest

GAS = (1.4, 0.0, 0.0, 2.5, 0.0, 1.0)
HEAVY = (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)
WATER = (4.4, 6.0e8, 0.0, 4160.0, 230.7, 1000.0)
AIR = (1.4, 0.0, 0.0, 718.0, 0.0, 1.25)

PAIRS = {
    "gas/heavy": ([GAS, HEAVY], dict(c=1.0, P=1.0, r=[1.0, 4.0])),
    "water/air": ([WATER, AIR], dict(c=30.0, P=1.0e5, r=[1000.0, 1.2])),
}


def _q_f(z, c):
    """Conserved vector and physical x-flux of cell c's state (9 components)."""
    r, u, v, P, e = z["rho"][c], z["u"][c], z["v"][c], z["P"][c], z["e"][c]
    b = np.zeros(4)
    b[:z["beta"].shape[0]] = z["beta"][:, c]
    E = e + 0.5 * (u * u + v * v)
    q = np.array([r, r * u, r * v, 0.0, r * E, *(r * b)])
    f = np.array([r * u, r * u * u + P, r * u * v, 0.0, (r * E + P) * u, *(r * u * b)])
    return q, f


def _mirror(F):
    G = F.copy()
    G[[0, 2, 3, 4, 5, 6, 7, 8]] *= -1.0
    return G


def translated_contact_error(oracle, pair, seed, n=600, cfl=0.9):
    """Largest relative deviation of the CRP sub-problem fluxes from the exact
    flux of a contact translated at uniform u and P."""
    mats, sc = PAIRS[pair]
    rng = np.random.default_rng(seed)
    m = oracle.Mesh.rectangle(3, 1, bounds=(2,) * 4)
    mm = oracle.MmFluid(m, mats, cfl=0.5)
    worst, count = 0.0, 0
    for _ in range(n):
        u = rng.uniform(-3.0, 3.0) * sc["c"]
        if abs(u) < 1e-3 * sc["c"]:
            continue
        v = rng.uniform(-0.5, 0.5) * sc["c"]
        P = rng.uniform(0.5, 2.0) * sc["P"]
        mat = rng.integers(0, 2, 3)
        dens = np.array([rng.uniform(0.5, 2.0) * sc["r"][k] for k in mat])
        inverse = bool(rng.random() < 0.5)
        if not inverse and u < 0:     # classic, clear to mix: no jump at the face
            mat[2], dens[2] = mat[1], dens[1]
        if inverse and u > 0:         # its mirror image
            mat[0], dens[0] = mat[1], dens[1]
        beta = np.zeros((2, 3))
        beta[mat, np.arange(3)] = 1.0
        mm.set_state(dens, np.full(3, u), np.full(3, v), np.full(3, P), beta)
        z = mm.state()
        g = np.array([mats[k][0] for k in mat])
        p0 = np.array([mats[k][1] for k in mat])
        c = np.sqrt(g * (P + p0) / dens)
        h = 0.01
        delta = rng.uniform(0.02, 1.0) * h
        dt = rng.uniform(0.02, 1.0) * cfl * h / (abs(u) + c.max())
        F, err = mm.crp_sub(inverse, 0, 1, 2, delta, dt)
        assert err == 0
        q, f = zip(*(_q_f(z, k) for k in range(3)))
        if not inverse:   # classic(LA, LB, RB): LB between the interface at -delta and the face
            want = (f[1] if dt <= delta / u else f[0] - delta * (q[0] - q[1]) / dt) if u > 0 else f[2]
        else:             # inverse(LA, RA, RB): RA between the face and the interface at +delta
            want = (f[1] if dt <= delta / -u else f[2] + delta * (q[2] - q[1]) / dt) if u < 0 else f[0]
        worst = max(worst, np.max(np.abs(F - want)) / np.max(np.abs(want)))
        count += 1
    assert count > n // 2
    return worst


def slow_fast_contact_error(oracle, seed, n=300):
    """The classic problem with a fast material A behind a slow, supersonic B
    (the wave pattern's first branch with the interface reaching the face):
    relative deviation from the exact translated-contact flux."""
    rng = np.random.default_rng(seed)
    m = oracle.Mesh.rectangle(3, 1, bounds=(2,) * 4)
    mm = oracle.MmFluid(m, [GAS, HEAVY], cfl=0.5)
    worst = 0.0
    for _ in range(n):
        dA, dB, dR = rng.uniform(0.8, 1.2), rng.uniform(3.5, 4.5), rng.uniform(3.5, 4.5)
        cA, cB = np.sqrt(1.4 / dA), np.sqrt(1.1 / dB)
        u = rng.uniform(1.05 * cB, 0.95 * cA)
        mm.set_state(np.array([dA, dB, dR]), np.full(3, u), np.full(3, 0.1), np.ones(3),
                     np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 1.0]]))
        z = mm.state()
        h = 0.01
        dt = rng.uniform(0.3, 1.0) * 0.9 * h / (u + max(cA, cB))
        delta = rng.uniform(0.01, 1.0) * dt * u
        F, err = mm.crp_sub(False, 0, 1, 2, delta, dt)
        assert err == 0
        q, f = zip(*(_q_f(z, k) for k in range(3)))
        want = f[1] if dt <= delta / u else f[0] - delta * (q[0] - q[1]) / dt
        worst = max(worst, np.max(np.abs(F - want)) / np.max(np.abs(want)))
    return worst


def mirror_error(oracle, pair, seed, n=300):
    """Largest relative deviation of calc_crp_flux(R', L') from the mirror
    image of calc_crp_flux(L, R), primes denoting reversed x-velocity."""
    mats, sc = PAIRS[pair]
    rng = np.random.default_rng(seed)
    m = oracle.Mesh.rectangle(4, 1, bounds=(2,) * 4)
    mm = oracle.MmFluid(m, mats, cfl=0.5)
    worst, count = 0.0, 0
    for _ in range(n):
        kind = rng.integers(0, 4)
        bL = rng.uniform(0.05, 0.95) if kind in (0, 1) else float(rng.integers(0, 2))
        bR = rng.uniform(0.05, 0.95) if kind in (0, 2) else float(rng.integers(0, 2))
        if bL in (0.0, 1.0) and bR in (0.0, 1.0):
            continue
        rA, rB = sc["r"]
        rL = 1.0 / (bL / rA + (1.0 - bL) / rB) * rng.uniform(0.8, 1.25)
        rR = 1.0 / (bR / rA + (1.0 - bR) / rB) * rng.uniform(0.8, 1.25)
        uL, uR = rng.uniform(-0.5, 0.5, 2) * sc["c"]
        vL, vR = rng.uniform(-0.3, 0.3, 2) * sc["c"]
        PL, PR = sc["P"] * rng.uniform(0.8, 1.25, 2)
        mm.set_state(np.array([rL, rR, rR, rL]), np.array([uL, uR, -uR, -uL]), np.array([vL, vR, vR, vL]),
                     np.array([PL, PR, PR, PL]), np.stack([np.array([bL, bR, bR, bL]), 1.0 - np.array([bL, bR, bR, bL])]))
        z = mm.state()
        alpha = [z["beta"][0][c] * z["rho"][c] / z["rhos"][0][c] if z["beta"][0][c] > 0 else 0.0 for c in (0, 1)]
        lo, hi = min(alpha), max(alpha)
        a_sig = float(np.clip(0.5 * (lo + hi) + 0.1 * rng.uniform(-1, 1), lo, hi))
        dt = rng.uniform(0.1, 1.0) * 0.5 * 0.01 / (3.0 * sc["c"])
        F1, e1 = mm.crp_flux(0, 1, 0.01, 0.01, 0, a_sig, dt)
        F2, e2 = mm.crp_flux(2, 3, 0.01, 0.01, 0, a_sig, dt)
        if e1 or e2:
            continue
        worst = max(worst, np.max(np.abs(F1 - _mirror(F2))) / np.max(np.abs(F1)))
        count += 1
    assert count > n // 2
    return worst


def split_error(oracle, pair, seed, n=200):
    """Mass, internal energy and closure consistency of PState::split."""
    mats, sc = PAIRS[pair]
    rng = np.random.default_rng(seed)
    m = oracle.Mesh.rectangle(1, 1, bounds=(2,) * 4)
    mm = oracle.MmFluid(m, mats + [GAS], cfl=0.5)
    worst = 0.0
    for _ in range(n):
        b = rng.dirichlet([1.0, 1.0, 1.0]) if rng.random() < 0.5 else np.array([*rng.dirichlet([1.0, 1.0]), 0.0])
        rA, rB = sc["r"]
        rho = 1.0 / (b[0] / rA + b[1] / rB + b[2] / 1.0) * rng.uniform(0.9, 1.1)
        mm.set_state(np.array([rho]), np.array([0.2]), np.array([-0.1]), np.array([sc["P"] * rng.uniform(0.5, 2)]),
                     b.reshape(3, 1))
        z = mm.state()
        present = z["beta"][:, 0] > 0.0
        if np.any(z["rhos"][:, 0][present] <= 0.0):
            continue   # below water's T0: no physical equilibrium (negative true density)
        iA = int(rng.integers(0, 3))
        if not (0.0 < z["beta"][iA][0] < 1.0):
            continue
        sp = mm.split(0, iA)
        alphaA = z["beta"][iA][0] * z["rho"][0] / z["rhos"][iA][0]
        bA = z["beta"][iA][0]
        mass = alphaA * sp["rhoA"] + (1.0 - alphaA) * sp["rhoB"]
        energy = bA * sp["eA"] + (1.0 - bA) * sp["eB"]
        worst = max(worst, abs(mass - z["rho"][0]) / z["rho"][0],
                    abs(energy - z["e"][0]) / abs(z["e"][0]))
        assert sp["P"] == z["P"][0] and sp["T"] == z["T"][0]
        # the mixture part closes back to the same (P, T)
        beta_B = sp["betaB"][:3]
        P2, T2, _ = oracle.mm_closure(mats + [GAS], sp["rhoB"], sp["eB"], beta_B, z["P"][0], z["T"][0])
        worst = max(worst, abs(P2 - z["P"][0]) / (abs(z["P"][0]) + sc["P"] * 1e-3), abs(T2 - z["T"][0]) / z["T"][0])
    return worst


def resting_interface_error(oracle, steps=20):
    """Water | mixed cell | air at rest in pressure equilibrium, with
    composite fluxes: largest |u| / c, relative pressure change and change
    of any cell's partial mass after `steps` steps."""
    m = oracle.Mesh.rectangle(8, 2, bounds=(1, 1, 3, 3))
    n = m.n
    cx = m.cells()["cx"]
    bw = np.where(cx < 0.45, 1.0, np.where(cx < 0.55, 0.5, 0.0))
    rho = 1.0 / (bw / 1000.0 + (1.0 - bw) / 1.2)
    mm = oracle.MmFluid(m, [WATER, AIR], cfl=0.5, crp=True)
    mm.set_state(rho, np.zeros(n), np.zeros(n), np.full(n, 1.0e5), np.stack([bw, 1.0 - bw]))
    z0 = mm.state()
    part0 = z0["rho"] * z0["beta"]
    for _ in range(steps):
        assert mm.update() == 0
    z = mm.state()
    c = np.sqrt(1.4 * 1.0e5 / 1.2)
    du = max(np.max(np.abs(z["u"])), np.max(np.abs(z["v"]))) / c
    dP = np.max(np.abs(z["P"] - 1.0e5)) / 1.0e5
    dm = np.max(np.abs(z["rho"] * z["beta"] - part0)) / np.max(part0)
    return max(du, dP, dm)


@pytest.mark.parametrize("pair,seed", [("gas/heavy", 3), ("water/air", 4)])
def test_crp_subproblems_translate_a_uniform_contact_exactly(oracle, pair, seed):
    assert translated_contact_error(oracle, pair, seed) < 1e-12


def test_crp_fast_behind_slow_supersonic(oracle):
    assert slow_fast_contact_error(oracle, 9) < 1e-13


@pytest.mark.parametrize("pair,seed", [("gas/heavy", 1), ("water/air", 2)])
def test_crp_flux_mirror_antisymmetric(oracle, pair, seed):
    assert mirror_error(oracle, pair, seed) < 1e-12


@pytest.mark.parametrize("pair,seed", [("gas/heavy", 5), ("water/air", 6)])
def test_split_keeps_mass_energy_and_equilibrium(oracle, pair, seed):
    assert split_error(oracle, pair, seed) < 1e-11


def test_resting_interface_stays_at_rest(oracle):
    assert resting_interface_error(oracle) < 1e-12


def test_planted_mistakes_are_caught(oracle):
    """Each planted mistake in the composite-flux chain breaks a check."""
    try:
        for k in (1, 2, 3, 4):
            oracle.plant(k)
            errs = [translated_contact_error(oracle, "gas/heavy", 3), translated_contact_error(oracle, "water/air", 4),
                    slow_fast_contact_error(oracle, 9), mirror_error(oracle, "gas/heavy", 1),
                    mirror_error(oracle, "water/air", 2), split_error(oracle, "gas/heavy", 5),
                    split_error(oracle, "water/air", 6)]
            assert max(errs) > 1e-6, (k, errs)
    finally:
        oracle.plant(0)
