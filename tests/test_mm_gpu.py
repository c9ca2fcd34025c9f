"""Device multi-material step (zfv_mm_*: MmFluid first order with and
without composite Riemann problems at mixed faces, PT-equilibrium closure
of stiffened gases, PLIC interface reconstruction) against the oracle on the
same inputs.

Bar: bit-exact (numpy equality, NaN where a material is absent) for the
state after the closure, every field and dt after steps on uniform and
adapted meshes, and the interface normals and positions; the closure's
Newton iterations take the same operations in the same order on both
sides.  One material must reproduce the device single-material first-order
step bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAS = (1.4, 0.0, 0.0, 2.5, 0.0, 1.0)
HEAVY = (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)
WATER = (4.4, 6.0e8, 0.0, 4160.0, 230.7, 1000.0)
AIR = (1.4, 0.0, 0.0, 718.0, 0.0, 1.25)
FIELDS = ("rho", "u", "v", "P", "e", "T", "beta", "rhos")


def initial(cx, cy, mats, seed, p_scale=1.0):
    """Smooth seeded fields; mass fractions varying across the domain (pure
    and mixed cells), the density from the fractions and the materials'
    reference densities (1 / sum beta_i / rho_i), perturbed by 10 %."""
    rng = np.random.default_rng(seed)
    n, nm = len(cx), len(mats)
    b = [np.clip(0.5 + 0.6 * np.sin((5 + k) * cx + (2 - k) * cy), 0.0, 1.0) for k in range(nm)]
    b = np.stack(b) + 1e-3 * (rng.random((nm, n)) < 0.05)
    b[0][b.sum(axis=0) == 0.0] = 1.0
    b = b / b.sum(axis=0)
    ref = np.array([m[5] for m in mats])
    r = (1.0 + 0.1 * np.sin(3 * cx) * np.cos(2 * cy)) / np.sum(b / ref[:, None], axis=0)
    u = 0.3 * np.cos(2 * cy) + 0.02 * rng.random(n)
    v = -0.2 * np.sin(3 * cx) + 0.02 * rng.random(n)
    p = p_scale * (1.0 + 0.3 * np.cos(4 * cx + cy))
    return r, u, v, p, b


def assert_same(zg, zo, what):
    for k in FIELDS:
        a, b = zg[k], zo[k]
        ok = np.array_equal(a, b, equal_nan=True)
        if not ok:
            bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))).reshape(-1))
            raise AssertionError(f"{what}: {k} differs at {bad[:5]}: {a.reshape(-1)[bad[:3]]} vs {b.reshape(-1)[bad[:3]]}")


def pair(oracle, zfv, nx, ny, bounds, mats, crp, cfl=0.5):
    m = oracle.Mesh.rectangle(nx, ny, bounds=bounds)
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, ny, bounds=bounds)
    ctx.mm_configure(mats, cfl, crp=crp)
    o = oracle.MmFluid(m, mats, cfl, crp=crp)
    return m, o, ctx


CASES = [
    ([GAS, HEAVY], (2, 1, 3, 3), 1.0),
    ([GAS, HEAVY, (1.67, 0.0, 0.0, 1.5, 0.0, 0.5)], (1, 1, 2, 2), 1.0),
    ([WATER, AIR], (2, 2, 1, 2), 1.0e5),
    ([GAS], (3, 3, 3, 3), 1.0),
]


@pytest.mark.parametrize("crp", [False, True])
@pytest.mark.parametrize("mats,bounds,sp", CASES)
def test_steps_match_oracle(oracle, zfv, mats, bounds, sp, crp):
    m, o, ctx = pair(oracle, zfv, 30, 22, bounds, mats, crp)
    cl = m.cells()
    st = initial(cl["cx"], cl["cy"], mats, 4, sp)
    o.set_state(*st)
    ctx.mm_set_state(*st)
    assert_same(ctx.mm_state(), o.state(), "after set_state")
    for step in range(12):
        assert o.update() == 0
        ctx.mm_step(1)
        assert ctx.mm_dt() == o.dt(), f"dt at step {step}"
        assert_same(ctx.mm_state(), o.state(), f"step {step + 1}")
    gx, gy, gp = ctx.mm_interface()
    ox, oy, op = o.interface()
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy) and np.array_equal(gp, op), "interface"
    if len(mats) > 1:
        assert np.any(gp != 0.0), "no mixed cell"


@pytest.mark.parametrize("crp", [False, True])
def test_steps_on_an_adapted_mesh(oracle, zfv, crp):
    """Hanging faces (second sub-face slots): refine both meshes with the same
    random flags through the single-material solvers, then run the
    multi-material step on the adapted meshes."""
    nx, ny, bounds = 20, 16, (2, 1, 2, 1)
    m = oracle.Mesh.rectangle(nx, ny, bounds=bounds)
    s = oracle.SmFluid(m, acc=1, flux=3)
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, ny, bounds=bounds)
    ctx.configure(acc=1, flux=3, limiter="MC", cfl=0.5, gamma=1.4)
    q = m.quadrant_state(3)
    s.set_state(*q)
    ctx.set_state(*q)
    rng = np.random.default_rng(9)
    for _ in range(2):
        f = rng.choice([-1, 0, 1], size=m.n, p=[0.1, 0.6, 0.3]).astype(np.int32)
        m.set_flags(f)
        ctx.set_flags_array(f)
        s.refine(2, slope=False)
        ctx.refine(2, slope=False)
    assert ctx.n == m.n
    _, slots, _ = ctx.mesh_download()
    assert np.any(slots["boundary"].reshape(-1, 8)[:, 4:] >= 0), "no sub-face slots"
    mats = [GAS, HEAVY]
    o = oracle.MmFluid(m, mats, 0.5, crp=crp)
    ctx.mm_configure(mats, 0.5, crp=crp)
    cl = m.cells()
    st = initial(cl["cx"], cl["cy"], mats, 6)
    o.set_state(*st)
    ctx.mm_set_state(*st)
    for step in range(8):
        assert o.update() == 0
        ctx.mm_step(1)
        assert ctx.mm_dt() == o.dt(), f"dt at step {step}"
        assert_same(ctx.mm_state(), o.state(), f"adapted, step {step + 1}")
    gx, gy, gp = ctx.mm_interface()
    ox, oy, op = o.interface()
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy) and np.array_equal(gp, op)


def test_one_material_equals_single_fluid_step(zfv):
    """The device multi-material step with one ideal gas and the device
    single-material first-order HLLC step agree bit for bit at equal dt."""
    from zephyr_amd import problems
    nx, ny, bounds = 40, 30, (1, 2, 3, 3)
    a = zfv.Context()
    a.mesh_rectangle(nx, ny, bounds=bounds)
    a.mm_configure([GAS], 0.5, crp=True)
    b = zfv.Context()
    b.mesh_rectangle(nx, ny, bounds=bounds)
    b.configure(acc=1, flux="hllc", limiter="MC", cfl=0.5, gamma=1.4)
    cells, _, _ = a.mesh_download()
    r, u, v, p, e = problems.four_quadrant(cells["cx"], cells["cy"], 12)
    a.mm_set_state(r, u, v, p, np.ones((1, a.n)))
    z = a.mm_state()
    b.set_state(z["rho"], z["u"], z["v"], z["P"], z["e"])
    a.mm_set_max_dt(1e-3)
    b.set_max_dt(1e-3)
    for step in range(10):
        a.mm_step(1)
        b.step(1)
        za, zb = a.mm_state(), b.state()
        for k, x in zip(("rho", "u", "v", "P", "e"), zb):
            assert np.array_equal(za[k], x), f"{k} at step {step}"


def test_state_is_tied_to_the_mesh(zfv):
    ctx = zfv.Context()
    ctx.mesh_rectangle(8, 8)
    ctx.mm_configure([GAS, HEAVY], 0.5)
    n = ctx.n
    ctx.mm_set_state(np.ones(n), np.zeros(n), np.zeros(n), np.ones(n), np.stack([np.ones(n), np.zeros(n)]))
    ctx.mm_step(2)
    ctx.mesh_rectangle(10, 8)
    with pytest.raises(zfv.ZfvError, match="no multi-material state"):
        ctx.mm_step(1)


def test_composite_fluxes_keep_an_advected_interface_sharp(zfv):
    """The device step with composite Riemann fluxes keeps an advected
    interface a few cells wide; plain HLLC smears it (as in the oracle)."""
    counts = {}
    for crp in (False, True):
        ctx = zfv.Context()
        ctx.mesh_rectangle(40, 40, bounds=(3, 3, 3, 3))
        ctx.mm_configure([GAS, HEAVY], 0.5, crp=crp)
        cells, _, _ = ctx.mesh_download()
        inside = ((cells["cx"] - 0.5) ** 2 + (cells["cy"] - 0.5) ** 2 < 0.2 ** 2).astype(float)
        n = ctx.n
        ctx.mm_set_state(np.where(inside > 0, 4.0, 1.0), np.full(n, 1.0), np.full(n, 0.5), np.ones(n),
                         np.stack([1.0 - inside, inside]))
        ctx.mm_step(60)
        a = ctx.mm_state()["beta"][1]
        counts[crp] = int(np.sum((a > 1e-6) & (a < 1.0 - 1e-6)))
    assert counts[True] < 0.2 * counts[False], counts
