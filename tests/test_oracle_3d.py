"""The oracle's three-dimensional SmFluid (oracle/oracle_3d.hpp) pinned by
what the mathematics fixes: the rotation to the face frame is orthogonal
and inverted by its global counterpart; a uniform state on a periodic box
stays uniform bit for bit; mass, momentum and energy are conserved on a
periodic box; the least-squares gradient of a linear field is exact away
from the walls; a problem uniform in z reproduces the two-dimensional
oracle; a one-dimensional shock tube gives the same profile along x, y and
z; and the Cuboid mesh matches its closed forms (volumes, centres, areas,
incircle, cell order c = nz (ny i + j) + k)."""
import numpy as np
import pytest

from oracle import pyoracle as O

GAMMA = 1.4


def energy(rho, p):
    return p / ((GAMMA - 1.0) * rho)


def test_rotation_orthogonal_and_inverse():
    rng = np.random.default_rng(1)
    normals = [np.array(n, float) for n in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))]
    normals += [v / np.linalg.norm(v) for v in rng.normal(size=(40, 3))]
    for n in normals:
        for _ in range(5):
            v = rng.normal(size=3)
            loc = O.rotate3(v, n, True)
            assert abs(loc[0] - np.dot(v, n)) < 1e-14, "local x is the normal component"
            assert abs(np.linalg.norm(loc) - np.linalg.norm(v)) < 1e-14
            np.testing.assert_allclose(O.rotate3(loc, n, False), v, rtol=0, atol=1e-14)


def test_cuboid_closed_forms():
    nx, ny, nz = 3, 4, 5
    m = O.Mesh3(nx, ny, nz, 0.0, 1.5, -1.0, 1.0, 0.0, 0.5, bounds=(O.WALL, O.ZOE, O.PERIODIC, O.PERIODIC, 2, 1))
    c, sl = m.cells(), m.slots()
    hx, hy, hz = 1.5 / nx, 2.0 / ny, 0.5 / nz
    i, j, k = np.divmod(np.arange(m.n) // nz, ny) + (np.arange(m.n) % nz,)
    np.testing.assert_allclose(c["cx"], (i + 0.5) * hx, rtol=1e-15)
    np.testing.assert_allclose(c["cy"], -1.0 + (j + 0.5) * hy, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(c["cz"], (k + 0.5) * hz, rtol=1e-15)
    assert np.all(c["volume"] == hx * hy * hz)
    np.testing.assert_allclose(c["h"], min(hx, hy, hz), rtol=1e-15)
    b = sl["boundary"].reshape(-1, 24)
    assert np.all(b[:, 6:] == O.UNDEFINED)
    assert np.all(b[i == 0, 0] == O.WALL) and np.all(b[i == nx - 1, 1] == O.ZOE)
    assert np.all(b[:, 2:4][(j == 0) | (j == ny - 1)] != O.WALL)   # periodic in y: both ends
    adj = sl["adj"].reshape(-1, 24)
    assert np.all(adj[(j == 0), 2] == (np.arange(m.n)[j == 0] + (ny - 1) * nz))
    area = sl["area"].reshape(-1, 24)
    assert np.all(area[:, 0] == hy * hz) and np.all(area[:, 4] == hx * hy)


def run_box(flux, acc, limiter, state, steps, shape=(6, 5, 4), bounds=(O.PERIODIC,) * 6):
    m = O.Mesh3(*shape, bounds=bounds)
    s = O.SmFluid3(m, flux=flux, acc=acc, limiter=limiter, cfl=0.5)
    s.set_state(*state(m))
    for _ in range(steps):
        assert s.update(2) == 0
    return m, s


@pytest.mark.parametrize("flux", [O.HLLC, O.HLL, O.RUSANOV])
@pytest.mark.parametrize("acc", [1, 2])
def test_uniform_state_stays_uniform(flux, acc):
    def state(m):
        n = m.n
        return (np.full(n, 1.3), np.full(n, 0.3), np.full(n, -0.2), np.full(n, 0.1), np.full(n, 2.0),
                np.full(n, energy(1.3, 2.0)))
    m, s = run_box(flux, acc, "MC", state, 1)
    first = s.state()
    for k, v in enumerate(state(m)[:4]):
        assert np.array_equal(first[k], v), f"component {k}"
    # p and e come back from E (round-off), then stay fixed bit for bit
    for _ in range(5):
        s.update(2)
    for k, (a, b) in enumerate(zip(s.state(), first)):
        assert np.array_equal(a, b), f"component {k} after 6 steps"
        assert np.all(a == a[0]), f"component {k} not uniform"


def blob(m, seed=4):
    rng = np.random.default_rng(seed)
    c = m.cells()
    r2 = (c["cx"] - 0.4) ** 2 + (c["cy"] - 0.55) ** 2 + (c["cz"] - 0.5) ** 2
    rho = 1.0 + 0.5 * np.exp(-r2 / 0.05)
    p = 1.0 + 2.0 * np.exp(-r2 / 0.03)
    u = 0.2 + 0.05 * rng.normal(size=m.n)
    v = -0.1 + 0.05 * rng.normal(size=m.n)
    w = 0.15 + 0.05 * rng.normal(size=m.n)
    return rho, u, v, w, p, energy(rho, p)


@pytest.mark.parametrize("acc,limiter", [(1, "MC"), (2, "MC"), (2, "minmod"), (2, "vanLeer")])
def test_conservation_on_periodic_box(acc, limiter):
    m = O.Mesh3(8, 7, 6, bounds=(O.PERIODIC,) * 6)
    s = O.SmFluid3(m, flux=O.HLLC, acc=acc, limiter=limiter)
    s.set_state(*blob(m))
    t0 = s.totals()
    for _ in range(10):
        assert s.update(2) == 0
    t1 = s.totals()
    scale = np.array([t0[0], t0[4], t0[4], t0[4], t0[4]])
    assert np.all(np.abs(t1 - t0) <= 1e-13 * scale), (t1 - t0) / scale


def test_lsm_gradient_of_a_linear_field_is_exact():
    m = O.Mesh3(7, 6, 5, 0.0, 1.0, 0.0, 1.2, 0.0, 0.8, bounds=(O.ZOE,) * 6)
    s = O.SmFluid3(m, flux=O.HLLC, acc=2, limiter="unit")
    c = m.cells()
    a = np.array([0.3, -0.7, 1.1])
    rho = 2.0 + a[0] * c["cx"] + a[1] * c["cy"] + a[2] * c["cz"]
    p = np.full(m.n, 1.0)
    s.set_state(rho, np.zeros(m.n), np.zeros(m.n), np.zeros(m.n), p, energy(rho, p))
    s.update(1)
    g = s.grad()   # the gradient of the state before the step
    nx, ny, nz = m.shape
    i, j, k = np.divmod(np.arange(m.n) // nz, ny) + (np.arange(m.n) % nz,)
    inner = (i > 0) & (i < nx - 1) & (j > 0) & (j < ny - 1) & (k > 0) & (k < nz - 1)
    for d in range(3):
        np.testing.assert_allclose(g[d, 0][inner], a[d], rtol=1e-12)
        assert np.all(np.abs(g[d, 1:4][:, inner]) < 1e-14)


@pytest.mark.parametrize("acc", [1, 2])
def test_uniform_in_z_reproduces_the_two_dimensional_oracle(acc):
    nx, ny, nz = 10, 8, 3
    m3 = O.Mesh3(nx, ny, nz, bounds=(O.WALL, O.ZOE, O.PERIODIC, O.PERIODIC, O.PERIODIC, O.PERIODIC))
    m2 = O.Mesh.rectangle(nx, ny, bounds=(O.WALL, O.ZOE, O.PERIODIC, O.PERIODIC))
    st2 = m2.quadrant_state(5)
    s2 = O.SmFluid(m2, flux=O.HLLC, acc=acc, limiter="MC")
    s2.set_state(*st2)
    rep = [np.repeat(a, nz) for a in st2]   # cell (i, j, k) of 3D = cell (i, j) of 2D
    s3 = O.SmFluid3(m3, flux=O.HLLC, acc=acc, limiter="MC")
    s3.set_state(rep[0], rep[1], rep[2], np.zeros(m3.n), rep[3], rep[4])
    for _ in range(6):
        s2.update(1)
        s3.update(1)
        assert abs(s3.dt() - s2.dt()) <= 1e-14 * s2.dt()
    r2, u2, v2, p2, e2 = s2.state()
    r3, u3, v3, w3, p3, e3 = s3.state()
    for a3, a2, name in ((r3, r2, "rho"), (u3, u2, "u"), (v3, v2, "v"), (p3, p2, "p")):
        np.testing.assert_allclose(a3.reshape(-1, nz), np.repeat(a2[:, None], nz, 1), rtol=1e-12, atol=1e-13,
                                   err_msg=name)
    assert np.all(np.abs(w3) < 1e-13)


@pytest.mark.parametrize("acc", [1, 2])
def test_shock_tube_is_the_same_along_each_axis(acc):
    n = 24
    res = []
    for axis in range(3):
        shape = [2, 2, 2]
        shape[axis] = n
        bounds = [O.PERIODIC] * 6
        bounds[2 * axis], bounds[2 * axis + 1] = O.WALL, O.ZOE
        m = O.Mesh3(*shape, bounds=tuple(bounds))
        c = m.cells()
        x = (c["cx"], c["cy"], c["cz"])[axis]
        left = x < 0.5
        rho = np.where(left, 1.0, 0.125)
        p = np.where(left, 1.0, 0.1)
        z = np.zeros(m.n)
        s = O.SmFluid3(m, flux=O.HLLC, acc=acc, limiter="minmod")
        s.set_state(rho, z, z, z, p, energy(rho, p))
        for _ in range(12):
            s.update(1)
        st = s.state()
        order = np.argsort(x, kind="stable")[::4]   # one cell per position along the axis
        res.append((st[0][order], st[1 + axis][order], st[4][order]))
    for axis in (1, 2):
        for a, b in zip(res[0], res[axis]):
            np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-13)
