"""The oracle's Transfer solver (math/solver/transfer.cpp) pinned by what the
mathematics fixes:

* compute_dt is CFL min(V / max S) / |v| (closed form on a uniform mesh);
* a uniform fraction in a uniform flow stays uniform bit for bit (every
  face flux is dt vn S a and they cancel);
* at Courant number 1 along x every CRP-bounded method (CRP_V3 / V5 / SE /
  N1, MUSCLd_CRP, MUSCL_MC_CRP) moves the profile exactly one cell: the
  bounds of flux_2D pinch to a V of the upwind cell;
* a linear profile in a uniform flow is advected exactly in the interior by
  MUSCL_MC and MUSCLd (exact LSM slopes, the MC limiter and the ATvL
  extrapolation both reproduce a linear field at the face), and the
  MUSCL_MC slopes are the field's gradient;
* the flux form conserves sum u V to round-off in a box with a cellular
  flow tangent to the walls for every method while u stays inside (0, 1) (each inner face's
  flux is the same from both sides);
* direction splitting: an x sweep moves nothing with a velocity along y;
* set_flags marks exactly the cells whose closed neighbourhood varies;
* WENO5 (weno.h) reproduces the face values of a quadratic from its cell
  averages (every sub-stencil is exact on quadratics), keeps a constant,
  takes the smooth side of a jump (non-oscillatory), and its m() is the
  mirror of p(); WENO and WENO_CRP are in the lists above (WENO_CRP is
  CRP-bounded, so Courant 1 moves the profile exactly one cell).
A swapped upwind side, a wrong bound, a wrong extrapolation offset or a
swapped interpolation weight fails one of these.
"""
import numpy as np
import pytest

ZOE, PERIODIC = 2, 3
BUILT = ["CRP_V3", "CRP_V5", "CRP_SE", "CRP_N1", "MUSCLd", "MUSCLd_CRP", "MUSCL_MC", "MUSCL_MC_CRP", "WENO",
         "WENO_CRP"]
BOUNDED = ["CRP_V3", "CRP_V5", "CRP_SE", "CRP_N1", "MUSCLd_CRP", "MUSCL_MC_CRP", "WENO_CRP"]


def velocity_field(m, f):
    cl, sl = m.cells(), m.slots()
    fvx, fvy = f(sl["fcx"], sl["fcy"])
    cvx, cvy = f(cl["cx"], cl["cy"])
    return fvx, fvy, cvx, cvy


def solver(oracle, m, method, f, cfl=0.5):
    T = oracle.Transfer(m, method, cfl)
    T.set_velocity(*velocity_field(m, f))
    return T


def uniform(vx, vy):
    return lambda x, y: (np.full_like(x, vx), np.full_like(y, vy))


def test_unbuilt_methods_are_refused(oracle):
    m = oracle.Mesh.rectangle(4, 4)
    for meth in ("CRP_N2", "VOF", "VOF_CRP", "MUSCLn", "MUSCLn_CRP"):
        with pytest.raises(ValueError):
            oracle.Transfer(m, meth)
    with pytest.raises(ValueError):   # five-cell stencils on a 2-cell-wide grid
        oracle.Transfer(oracle.Mesh.rectangle(2, 4), "WENO")


def test_compute_dt_closed_form(oracle):
    m = oracle.Mesh.rectangle(20, 10, 0.0, 2.0, 0.0, 0.5)   # hx = 0.1, hy = 0.05
    T = solver(oracle, m, "CRP_V3", uniform(0.6, 0.8), cfl=0.4)
    assert T.compute_dt() == pytest.approx(0.4 * (0.1 * 0.05 / 0.1) / 1.0, rel=1e-15)


@pytest.mark.parametrize("method", BUILT)
def test_uniform_state_stays_uniform(oracle, method):
    m = oracle.Mesh.rectangle(12, 10, bounds=(PERIODIC,) * 4)
    T = solver(oracle, m, method, uniform(0.7, -0.4))
    T.set_state(np.full(m.n, 0.35))
    T.set_dt(T.compute_dt())
    for d in ("ANY", "X", "Y"):
        T.update(d)
    u = T.state()[0]
    assert np.all(u == 0.35)


def stripe(m):
    cx = m.cells()["cx"]
    return np.clip(np.sin(2 * np.pi * cx) * 1.5, 0.0, 1.0) * (0.2 + 0.8 * (m.cells()["cy"] > 0.4))


@pytest.mark.parametrize("method", BOUNDED)
def test_courant_one_shifts_one_cell(oracle, method):
    nx, ny = 20, 8
    m = oracle.Mesh.rectangle(nx, ny, bounds=(PERIODIC,) * 4)
    T = solver(oracle, m, method, uniform(1.0, 0.0))
    u0 = stripe(m)
    T.set_state(u0)
    T.set_dt(1.0 / nx)   # gamma = dt |v| S = V exactly
    T.update("X")
    shifted = u0.reshape(nx, ny)[np.arange(nx) - 1].reshape(-1)   # cell (i, j) <- (i - 1, j)
    np.testing.assert_allclose(T.state()[0], shifted, rtol=0, atol=2e-15)


@pytest.mark.parametrize("method", ["MUSCL_MC", "MUSCLd"])
def test_linear_profile_is_advected_exactly(oracle, method):
    nx, ny = 24, 6
    m = oracle.Mesh.rectangle(nx, ny, bounds=(ZOE,) * 4)
    T = solver(oracle, m, method, uniform(0.8, 0.0))
    cx = m.cells()["cx"]
    b = 0.5
    T.set_state(0.2 + b * cx)
    dt = 0.3 / nx
    T.set_dt(dt)
    T.update("ANY")
    u = T.state()[0].reshape(nx, ny)
    want = (0.2 + b * cx - b * 0.8 * dt).reshape(nx, ny)
    np.testing.assert_allclose(u[3:-3], want[3:-3], rtol=0, atol=1e-14)
    if method == "MUSCL_MC":
        dx, dy = T.slopes()
        np.testing.assert_allclose(dx.reshape(nx, ny)[2:-2], b, rtol=1e-13)
        np.testing.assert_allclose(dy.reshape(nx, ny)[2:-2], 0.0, atol=1e-13)


@pytest.mark.parametrize("method", BUILT)
def test_rotation_conserves_volume(oracle, method):
    # boundary faces carry nothing in Transfer (fluxes skip them); across a
    # periodic face the reference extrapolates with the neighbour's own
    # centre on the far side, so its two sides differ: ZOE sides here
    n = 32
    m = oracle.Mesh.rectangle(n, n, bounds=(ZOE,) * 4)
    # the cellular flow of psi = sin(pi x) sin(pi y), tangent to the walls
    T = solver(oracle, m, method, lambda x, y: (np.pi * np.sin(np.pi * x) * np.cos(np.pi * y),
                                                -np.pi * np.cos(np.pi * x) * np.sin(np.pi * y)))
    cl = m.cells()
    r2 = (cl["cx"] - 0.6) ** 2 + (cl["cy"] - 0.5) ** 2
    u0 = 0.25 + 0.5 * np.exp(-r2 / 0.03)   # far from 0 and 1: no clamping
    T.set_state(u0)
    V = cl["volume"]
    total0 = np.sum(u0 * V)
    T.set_dt(T.compute_dt())
    for _ in range(10):
        T.update("ANY")
    u = T.state()[0]
    assert u.min() > 0.0 and u.max() < 1.0
    assert abs(np.sum(u * V) - total0) <= 1e-13 * total0


def test_x_sweep_ignores_y_velocity(oracle):
    m = oracle.Mesh.rectangle(10, 10, bounds=(PERIODIC,) * 4)
    T = solver(oracle, m, "CRP_V3", uniform(0.0, 1.0))
    u0 = np.linspace(0.1, 0.9, m.n)
    T.set_state(u0)
    T.set_dt(0.05)
    T.update("X")
    assert np.array_equal(T.state()[0], u0)
    T.update("Y")
    assert not np.array_equal(T.state()[0], u0)


def test_set_flags(oracle):
    nx, ny = 8, 6
    m = oracle.Mesh.rectangle(nx, ny, bounds=(ZOE,) * 4)
    T = solver(oracle, m, "CRP_V3", uniform(1.0, 0.0))
    u = np.zeros((nx, ny))
    u[3, 2] = 0.5
    T.set_state(u.reshape(-1))
    T.set_flags()
    flag = m.cells()["flag"].reshape(nx, ny)
    want = -np.ones((nx, ny), int)
    for i, j in ((3, 2), (2, 2), (4, 2), (3, 1), (3, 3)):
        want[i, j] = 1
    assert np.array_equal(flag, want)


def test_weno5_exact_on_quadratics(oracle):
    rng = np.random.default_rng(7)
    k = np.arange(-2, 3, dtype=float)
    for _ in range(50):
        a, b, c = rng.normal(size=3)
        avg = a + b * k + c * (k * k + 1.0 / 12.0)   # averages over [k - 1/2, k + 1/2]
        q = lambda x: a + b * x + c * x * x
        scale = abs(a) + abs(b) + abs(c)
        assert abs(oracle.weno5(avg, True) - q(0.5)) < 1e-13 * scale
        assert abs(oracle.weno5(avg, False) - q(-0.5)) < 1e-13 * scale


def test_weno5_constant_jump_and_mirror(oracle):
    for v in (0.0, 1.0, 0.3, 0.7):
        assert abs(oracle.weno5([v] * 5, True) - v) <= 4e-16   # a few ulps: the weights sum to 1 in rounding
    step = np.array([0.0, 0.0, 1.0, 1.0, 1.0])   # jump on the left face of the centre cell
    assert abs(oracle.weno5(step, True) - 1.0) < 1e-6
    assert abs(oracle.weno5(step, False) - 1.0) < 1e-6
    assert abs(oracle.weno5(step[::-1], True) - 1.0) < 1e-6
    rng = np.random.default_rng(3)
    for _ in range(50):
        u = rng.random(5)
        assert oracle.weno5(u[::-1], False) == pytest.approx(oracle.weno5(u, True), rel=1e-13, abs=1e-15)
