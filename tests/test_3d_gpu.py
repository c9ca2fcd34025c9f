"""Parity of the three-dimensional step (zfv3_* through the C ABI) with the
oracle's SmFluid on Cuboid meshes (oracle/oracle_3d.hpp).

Bar: bit-exact (value equality) for the mesh (centres, volumes, incircle
diameters, face centres, areas, least-squares weights), the CFL time step,
the state after first-order and MUSCL steps with the Rusanov, HLL and HLLC
fluxes and every limiter, the limited gradients and the isolation counters
-- both sides are the same sequence of correctly rounded IEEE operations.
HLLC_LM's sin and Godunov's pow are the device's (within an ulp of
glibc's): relative 1e-12 after four steps.  The device's conserved totals are a tree sum: relative
1e-13."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4
FLUX = {"godunov": 0, "rusanov": 1, "hll": 2, "hllc": 3, "hllc_m": 4, "hllc_lm": 5, "cir1": 6, "cir2": 7}
LIMITERS = ["none", "unit", "minmod", "MC", "vanLeer", "vanAlbada1", "vanAlbada2", "CHARM", "HCUS",
            "HQUICK", "Koren", "smart", "superbee"]   # ids 0..12 in both libraries


def blast_state(cx, cy, cz, seed=0, centre=(0.45, 0.5, 0.55)):
    """A spherical pressure and density jump with a random smooth velocity field."""
    rng = np.random.default_rng(seed)
    r = np.sqrt((cx - centre[0]) ** 2 + (cy - centre[1]) ** 2 + (cz - centre[2]) ** 2)
    a = rng.normal(size=(3, 3))
    rho = 1.0 + 0.5 * (r < 0.3) + 0.05 * np.sin(7.0 * cx + 3.0 * cy)
    u = 0.2 * np.sin(a[0, 0] * cx + a[0, 1] * cy + a[0, 2] * cz)
    v = 0.2 * np.cos(a[1, 0] * cx + a[1, 1] * cy + a[1, 2] * cz)
    w = 0.2 * np.sin(a[2, 0] * cx - a[2, 1] * cy + a[2, 2] * cz)
    p = 1.0 + 9.0 * (r < 0.25) + 0.1 * np.cos(5.0 * cz)
    e = p / ((GAMMA - 1.0) * rho)
    return rho, u, v, w, p, e


def make_pair(oracle, zfv, shape, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), bounds=(2,) * 6, acc=1, flux="hllc",
              limiter="MC", cfl=0.5, seed=0):
    m = oracle.Mesh3(*shape, *box, bounds=bounds)
    s = oracle.SmFluid3(m, flux=FLUX[flux], acc=acc, limiter=limiter, cfl=cfl)
    ctx = zfv.Context3()
    ctx.mesh_cuboid(*shape, *box, bounds=bounds)
    ctx.configure(acc=acc, flux=flux, limiter=limiter, cfl=cfl, gamma=GAMMA)
    c = m.cells()
    z = blast_state(c["cx"], c["cy"], c["cz"], seed)
    s.set_state(*z)
    ctx.set_state(*z)
    return m, s, ctx


def assert_equal(a, b, names, what):
    for k, name in enumerate(names):
        bad = np.flatnonzero(a[k] != b[k])
        assert bad.size == 0, f"{what}: {name} differs at {bad[:5]}: {a[k][bad[:3]]} vs {b[k][bad[:3]]}"


STATE = ["rho", "u", "v", "w", "p", "e"]


@pytest.mark.parametrize("shape,box,bounds", [
    ((3, 4, 5), (0.0, 1.5, -1.0, 1.0, 0.0, 0.5), (1, 2, 3, 3, 2, 1)),
    ((7, 1, 6), (-0.3, 1.7, 0.1, 0.35, 2.0, 2.9), (3, 3, 1, 1, 4, 2)),
    ((16, 16, 16), (0.0, 1.0, 0.0, 1.0, 0.0, 1.0), (2,) * 6),
])
def test_cuboid_mesh_matches_oracle(oracle, zfv, shape, box, bounds):
    m = oracle.Mesh3(*shape, *box, bounds=bounds)
    ctx = zfv.Context3()
    ctx.mesh_cuboid(*shape, *box, bounds=bounds)
    cells, slots = ctx.mesh()
    oc, osl = m.cells(), m.slots()
    for k, ok in (("cx", "cx"), ("cy", "cy"), ("cz", "cz"), ("vol", "volume"), ("h", "h")):
        assert np.array_equal(cells[k], oc[ok]), k
    for k in ("boundary", "adj", "nx", "ny", "nz", "fcx", "fcy", "fcz", "area", "w"):
        assert np.array_equal(slots[k], osl[k]), k


@pytest.mark.parametrize("flux", ["rusanov", "hll", "hllc"])
@pytest.mark.parametrize("acc", [1, 2])
@pytest.mark.parametrize("bounds", [(2,) * 6, (3,) * 6, (1, 1, 1, 1, 1, 1), (1, 2, 3, 3, 2, 1)])
def test_step_matches_oracle(oracle, zfv, flux, acc, bounds):
    m, s, ctx = make_pair(oracle, zfv, (9, 8, 7), bounds=bounds, acc=acc, flux=flux, seed=acc)
    assert ctx.compute_dt() == s.compute_dt()
    for step in range(4):
        assert s.update() == 0
        ctx.step(1)
        assert ctx.dt() == s.dt(), f"dt at step {step}"
        assert_equal(ctx.state(), s.state(), STATE, f"{flux} acc {acc} step {step}")
        if acc == 2:
            assert np.array_equal(ctx.grad(), s.grad()), f"gradient at step {step}"
            assert np.array_equal(ctx.wait(), s.wait())


@pytest.mark.parametrize("flux", ["hllc_m", "hllc_lm", "cir1", "cir2", "godunov"])
@pytest.mark.parametrize("acc", [1, 2])
def test_other_fluxes_match_oracle(oracle, zfv, flux, acc):
    m, s, ctx = make_pair(oracle, zfv, (9, 8, 7), bounds=(1, 2, 3, 3, 2, 1), acc=acc, flux=flux, seed=3)
    for step in range(4):
        assert s.update() == 0
        ctx.step(1)
    if flux not in ("hllc_lm", "godunov"):
        assert ctx.dt() == s.dt()
        assert_equal(ctx.state(), s.state(), STATE, flux)
    else:
        for a, b in zip(ctx.state(), s.state()):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b)))


@pytest.mark.parametrize("limiter", LIMITERS)
def test_muscl_limiters_match_oracle(oracle, zfv, limiter):
    m, s, ctx = make_pair(oracle, zfv, (6, 7, 8), box=(0.0, 1.0, 0.0, 1.2, -0.1, 0.9), bounds=(1, 2, 2, 1, 3, 3),
                          acc=2, flux="hllc", limiter=LIMITERS.index(limiter), seed=5)
    for step in range(3):
        s.update()
        ctx.step(1)
    assert_equal(ctx.state(), s.state(), STATE, limiter)
    assert np.array_equal(ctx.grad(), s.grad())


def test_isolated_cells_take_the_robust_flux(oracle, zfv):
    """Cells with wait > 0 keep their state through the predictor and their
    faces take HLL on the half-step states (sm_fluid.cpp:186-190, 280-290)."""
    m, s, ctx = make_pair(oracle, zfv, (8, 8, 8), bounds=(1,) * 6, acc=2, flux="hllc", seed=9)
    rng = np.random.default_rng(3)
    wait = np.where(rng.random(m.n) < 0.2, rng.integers(1, 4, m.n), 0).astype(np.int32)
    s.set_wait(wait)
    ctx.set_wait(wait)
    for step in range(4):
        s.update()
        ctx.step(1)
        assert_equal(ctx.state(), s.state(), STATE, f"step {step}")
        assert np.array_equal(ctx.wait(), s.wait())


def test_multi_step_launch_matches_oracle(oracle, zfv):
    """Several steps in one zfv3_smf_step call (the CFL slots alternate)."""
    m, s, ctx = make_pair(oracle, zfv, (10, 9, 11), bounds=(3, 3, 1, 1, 2, 2), acc=2, flux="hllc", seed=4)
    for _ in range(5):
        s.update()
    ctx.step(5)
    assert ctx.dt() == s.dt()
    assert_equal(ctx.state(), s.state(), STATE, "five steps")


def test_large_cube_matches_oracle(oracle, zfv):
    m, s, ctx = make_pair(oracle, zfv, (48, 48, 48), bounds=(2,) * 6, acc=2, flux="hllc", seed=11)
    for _ in range(2):
        s.update()
    ctx.step(2)
    assert_equal(ctx.state(), s.state(), STATE, "48^3")


def test_periodic_box_conserves(zfv):
    ctx = zfv.Context3()
    ctx.mesh_cuboid(24, 20, 16, bounds=(3,) * 6)
    ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
    cells, _ = ctx.mesh()
    ctx.set_state(*blast_state(cells["cx"], cells["cy"], cells["cz"], 2))
    t0 = ctx.totals()
    ctx.step(20)
    t1 = ctx.totals()
    scale = np.array([t0[0], t0[4], t0[4], t0[4], t0[4]])
    assert np.all(np.abs(t1 - t0) <= 1e-13 * np.abs(scale) * 20), (t0, t1)


def test_uniform_state_stays_uniform(zfv):
    ctx = zfv.Context3()
    ctx.mesh_cuboid(6, 5, 4, bounds=(1, 1, 2, 2, 3, 3))
    ctx.configure(acc=2, flux="hll", limiter="minmod", gamma=GAMMA)
    n = ctx.n
    ctx.set_state(np.full(n, 1.3), np.zeros(n), np.zeros(n), np.zeros(n), np.full(n, 2.0))
    ctx.step(3)
    rho, u, v, w, p, e = ctx.state()
    assert np.all(rho == 1.3) and np.all(p == 2.0)
    assert np.all(u == 0.0) and np.all(v == 0.0) and np.all(w == 0.0)


def test_errors(zfv):
    ctx = zfv.Context3()
    with pytest.raises(zfv.ZfvError):
        ctx.step(1)   # no mesh
    with pytest.raises(zfv.ZfvError):
        ctx.mesh_cuboid(0, 4, 4)
    ctx.mesh_cuboid(4, 4, 4)
    with pytest.raises(zfv.ZfvError) as ei:
        ctx.configure(flux=8)   # CRP is MmFluid's
    assert ei.value.code == zfv.ZFV_ERR_UNSUPPORTED
    ctx.configure(flux="hllc")
    with pytest.raises(zfv.ZfvError):
        ctx.step(1)   # no state


def test_kernel_timing(zfv):
    ctx = zfv.Context3()
    ctx.mesh_cuboid(16, 16, 16)
    ctx.configure(acc=2, flux="hllc", limiter="MC", gamma=GAMMA)
    n = ctx.n
    ctx.set_state(np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n))
    ctx.profile_enable(True)
    ctx.step(3)
    prof = ctx.profile_read()
    # the CFL reduction runs alone once after a new state, then inside the corrector
    assert {k: v[1] for k, v in prof.items()} == {"dt": 1, "grad_pred": 3, "corr": 3}
    assert all(v[0] > 0.0 for v in prof.values())
    ctx.profile_enable(False)


@pytest.mark.parametrize("bounds", [(1, 2, 3, 3, 2, 1), (3,) * 6])
@pytest.mark.parametrize("acc", [1, 2])
def test_geometry_tables_and_slot_arrays_agree(oracle, zfv, bounds, acc):
    """The step kernels read a Cuboid's 1D geometry tables (the default) or
    the per-slot arrays; both give the oracle's states."""
    m = oracle.Mesh3(7, 6, 9, bounds=bounds)
    s = oracle.SmFluid3(m, flux=3, acc=acc, limiter="MC", cfl=0.5)
    c = m.cells()
    z = blast_state(c["cx"], c["cy"], c["cz"], 8)
    s.set_state(*z)
    ctxs = []
    for on in (True, False):
        ctx = zfv.Context3()
        ctx.set_geometry_tables(on)
        ctx.mesh_cuboid(7, 6, 9, bounds=bounds)
        assert ctx.geometry_tables() == on
        ctx.configure(acc=acc, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
        ctx.set_state(*z)
        ctxs.append(ctx)
    for step in range(3):
        s.update()
        for ctx in ctxs:
            ctx.step(1)
            assert ctx.dt() == s.dt()
            assert_equal(ctx.state(), s.state(), STATE, f"tables {ctx.geometry_tables()} step {step}")
