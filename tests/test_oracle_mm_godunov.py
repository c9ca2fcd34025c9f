"""Known answers for the oracle's multi-material Godunov flux
(math/cfd/flux/godunov.cpp:39-77 with MixturePT::stiffened_gas,
mixture_pt.cpp:152-169), fixed by the mathematics rather than by the
formulas themselves:

* the tangent stiffened gas of a mixed state has the mixture's sound
  speed, c^2 = gamma (P + P0) / rho = (P / rho^2 - e_rho) / e_P (a wrong
  sign or factor in gamma or P0 breaks the identity); a pure state's is its
  own material;
* air driven into water: the face flux is the physical flux of the left
  star state, whose pressure solves f_L(P) + f_R(P) + u_R - u_L = 0 for two
  different stiffened gases (solved here with scipy's brentq on Toro's
  functions), density from the Rankine-Hugoniot shock relation, energy from
  the air's EOS;
* consistency F(z, z) = Flux(z) for mixed states, and supersonic flow takes
  the upwind state;
* with one ideal gas the multi-material Godunov step equals the
  single-material Godunov step bit for bit.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

WATER = (4.4, 6.0e8, 0.0, 4160.0, 230.7, 1000.0)
AIR = (1.4, 0.0, 0.0, 718.0, 0.0, 1.25)
GAS = (1.4, 0.0, 0.0, 2.5, 0.0, 1.0)
HEAVY = (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)


def two_cells(oracle, mats, left, right):
    """A 2 x 1 mesh holding (rho, u, v, P, beta) left and right."""
    m = oracle.Mesh.rectangle(2, 1)
    mm = oracle.MmFluid(m, mats, 0.5, crp=False)
    st = [np.array([left[k], right[k]]) for k in range(4)]
    beta = np.array([left[4], right[4]]).T
    mm.set_state(*st, beta)
    return m, mm


def phys_flux(z, c):
    ru = z["rho"][c] * z["u"][c]
    E = z["e"][c] + 0.5 * (z["u"][c] ** 2 + z["v"][c] ** 2)
    out = [ru, ru * z["u"][c] + z["P"][c], ru * z["v"][c], 0.0, (z["rho"][c] * E + z["P"][c]) * z["u"][c]]
    return np.array(out + [ru * b for b in z["beta"][:, c]] + [0.0] * (4 - z["beta"].shape[0]))


@pytest.mark.parametrize("b0", [0.3, 0.7, 0.95])
def test_tangent_gas_has_the_mixture_sound_speed(oracle, b0):
    for mats, P in (([GAS, HEAVY], 1.3), ([WATER, AIR], 2.0e5)):
        rho = 1.0 if mats[0] is GAS else 400.0
        _, mm = two_cells(oracle, mats, (rho, 0.1, 0.0, P, [b0, 1 - b0]), (rho, 0.0, 0.0, P, [1.0, 0.0]))
        g, p0 = mm.stiffened_gas(0)
        c = mm.sound_speed(0)
        assert g > 1.0
        assert abs(np.sqrt(g * (P + p0) / rho) - c) <= 1e-10 * c
        assert tuple(mm.stiffened_gas(1)) == (mats[0][0], mats[0][1]), "a pure state is its material"


def stiffened_f(P, rho, p, g, p0):
    c = np.sqrt(g * (p + p0) / rho)
    if P >= p:
        A, B = 2.0 / ((g + 1) * rho), (g - 1) / (g + 1) * p + 2 * g * p0 / (g + 1)
        return (P - p) * np.sqrt(A / (P + B))
    return 2 * c / (g - 1) * (((P + p0) / (p + p0)) ** ((g - 1) / (2 * g)) - 1)


def test_air_into_water_is_the_left_star_state(oracle):
    L = (1.25, 300.0, 5.0, 1.0e5, [0.0, 1.0])    # air, moving right
    R = (1000.0, 0.0, -2.0, 1.0e5, [1.0, 0.0])   # water at rest
    _, mm = two_cells(oracle, [WATER, AIR], L, R)
    F, err = mm.face_flux_kind("Godunov", 0, 1)
    assert err == 0
    gL, p0L, gR, p0R = AIR[0], AIR[1], WATER[0], WATER[1]
    f = lambda P: stiffened_f(P, L[0], L[3], gL, p0L) + stiffened_f(P, R[0], R[3], gR, p0R) + R[1] - L[1]
    P = brentq(f, L[3], 1.0e9, xtol=1e-14 * 1e9, rtol=1e-15)
    U = 0.5 * (L[1] + R[1]) + 0.5 * (stiffened_f(P, R[0], R[3], gR, p0R) - stiffened_f(P, L[0], L[3], gL, p0L))
    assert U > 0.0 and P > L[3]
    r = (P + p0L) / (L[3] + p0L)
    k = (gL - 1) / (gL + 1)
    rho = L[0] * (r + k) / (k * r + 1)          # Rankine-Hugoniot density behind the left shock
    D = L[1] + (P - L[3]) / (L[0] * (U - L[1]))
    assert D < 0.0, "the left shock moves left: the face is in the left star region"
    e = (P + gL * p0L) / ((gL - 1) * rho)
    want = np.array([rho * U, rho * U * U + P, rho * U * L[2], 0.0, (rho * (e + 0.5 * (U * U + L[2] ** 2)) + P) * U,
                     0.0, rho * U, 0.0, 0.0])
    np.testing.assert_allclose(F, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_consistency_and_supersonic(oracle):
    for mats in ([GAS, HEAVY], [WATER, AIR]):
        rho, P = (1.0, 1.0) if mats[0] is GAS else (500.0, 3.0e5)
        _, mm = two_cells(oracle, mats, (rho, 0.2, 0.1, P, [0.4, 0.6]), (rho, 0.2, 0.1, P, [0.4, 0.6]))
        z = mm.state()
        F, err = mm.face_flux_kind("Godunov", 0, 0)
        assert err == 0
        want = phys_flux(z, 0)
        np.testing.assert_allclose(F, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
    # supersonic to the right: the face carries the left state
    _, mm = two_cells(oracle, [GAS, HEAVY], (1.0, 5.0, 0.3, 1.0, [0.5, 0.5]), (0.8, 4.5, -0.1, 0.9, [1.0, 0.0]))
    z = mm.state()
    F, _ = mm.face_flux_kind("Godunov", 0, 1)
    want = phys_flux(z, 0)
    np.testing.assert_allclose(F, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_one_gas_equals_single_fluid_godunov(oracle):
    m = oracle.Mesh.rectangle(20, 16, bounds=(1, 2, 3, 3))
    cl = m.cells()
    rng = np.random.default_rng(3)
    n = m.n
    r = 1.0 + 0.3 * np.sin(3 * cl["cx"]) + 0.05 * rng.random(n)
    u = 0.4 * np.cos(2 * cl["cy"]) + 0.01
    v = -0.2 * np.sin(3 * cl["cx"]) + 0.013
    p = 1.0 + 0.4 * np.cos(4 * cl["cx"] + cl["cy"])
    mm = oracle.MmFluid(m, [GAS], cfl=0.5, crp=False)
    mm.set_method("Godunov")
    mm.set_state(r, u, v, p, np.ones((1, n)))
    sm = oracle.SmFluid(m, gamma=1.4, flux=oracle.GODUNOV, acc=1, cfl=0.5)
    z = mm.state()
    sm.set_state(z["rho"], z["u"], z["v"], z["P"], z["e"])
    for st in (mm, sm):
        st.set_max_dt(1e-3)
    for step in range(5):
        assert mm.update() == 0
        sm.update(nthreads=1)
        zm, zs = mm.state(), sm.state()
        for k, a in zip(("rho", "u", "v", "P", "e"), zs):
            assert np.array_equal(zm[k], a), f"{k} at step {step}"
