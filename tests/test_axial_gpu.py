"""Axisymmetric SmFluid on the device (zfv_mesh_rectangle_axial, params
axial = 1; Rectangle::set_axial, rectangle.cpp:97-103 and 590-602; SmFluid
m_axial, sm_fluid.cpp:160-169, 236-245, 341-350) against the oracle's
axisymmetric restatement (pinned in test_oracle_axial.py).

Bar: bit-exact geometry (centres, volume_alt, area_alt, boundaries) and
bit-exact state and dt after steps at both orders, with several fluxes and
limiters, in every launch layout of the split step (the tiled step, which
has no axisymmetric terms, is not taken with set_axial: the default
Rectangle launch would otherwise miss the oracle); the resting fluid stays
at rest on the device; the totals are the volume_alt-weighted ones;
adaptation and an axial solver on a planar mesh are refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4
WALL, ZOE, PERIODIC = 1, 2, 3


def pair(oracle, zfv, nx, ny, acc, flux, limiter="MC", ymin=0.0, bounds=(ZOE, ZOE, ZOE, WALL)):
    m = oracle.Mesh.rectangle(nx, ny, 0.0, 2.0, ymin, 1.0, bounds=bounds, axial=True)
    s = oracle.SmFluid(m, acc=acc, flux=flux, limiter=limiter, cfl=0.45)
    s.set_axial(True)
    ctx = zfv.Context()
    ctx.mesh_rectangle_axial(nx, ny, 0.0, 2.0, ymin, 1.0, bounds=bounds)
    ctx.configure(acc=acc, flux=flux, limiter=limiter, cfl=0.45, gamma=GAMMA, axial=True)
    return m, s, ctx


def test_geometry_matches_oracle(oracle, zfv):
    m, _, ctx = pair(oracle, zfv, 9, 7, 1, 3, ymin=0.0)
    cells, slots, verts = ctx.mesh_download()
    oc, osl = m.cells(), m.slots()
    for k in ("cx", "cy", "volume"):
        assert np.array_equal(cells[k], oc[k]), k
    for k in ("boundary", "adj", "fcx", "fcy", "area", "nx", "ny"):
        assert np.array_equal(slots[k], osl[k]), k
    va, aa = ctx.mesh_axial()
    ova, oaa = m.axial_geometry()
    assert np.array_equal(va, ova) and np.array_equal(aa, oaa)
    assert np.all(slots["boundary"].reshape(-1, 8)[np.arange(9) * 7, 2] == WALL)


@pytest.mark.parametrize("acc,flux,limiter", [(1, 3, "MC"), (1, 2, "MC"), (2, 3, "MC"), (2, 3, "minmod"),
                                              (2, 1, "vanLeer"), (2, 6, "MC")])
def test_steps_match_oracle(oracle, zfv, acc, flux, limiter):
    m, s, ctx = pair(oracle, zfv, 24, 18, acc, flux, limiter, bounds=(ZOE, WALL, ZOE, ZOE))
    st = m.quadrant_state(11)
    s.set_state(*st)
    ctx.set_state(*st)
    for step in range(10):
        assert s.update(2) == 0
        ctx.step(1)
        assert ctx.dt() == s.dt(), f"dt at step {step}"
        for a, b, name in zip(ctx.state(), s.state(), "rho u v p e".split()):
            assert np.array_equal(a, b), f"step {step + 1}: {name}"


@pytest.mark.parametrize("layout", [0, 1, 2])
def test_layouts_match_oracle(oracle, zfv, layout):
    m, s, ctx = pair(oracle, zfv, 32, 20, 2, 3, "MC", ymin=0.1, bounds=(PERIODIC, PERIODIC, WALL, WALL))
    for stage in (1, 2):
        ctx.set_launch(stage, layout, 256)
    st = m.quadrant_state(4)
    s.set_state(*st)
    ctx.set_state(*st)
    for step in range(6):
        assert s.update(2) == 0
        ctx.step(1)
    for a, b in zip(ctx.state(), s.state()):
        assert np.array_equal(a, b)
    np.testing.assert_allclose(ctx.totals(), s.totals(), rtol=1e-13, atol=1e-13)


def test_rest_stays_at_rest_and_refusals(zfv):
    ctx = zfv.Context()
    ctx.mesh_rectangle_axial(16, 12, 0.0, 1.0, 0.0, 1.0, bounds=(PERIODIC, PERIODIC, WALL, WALL))
    ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.4, gamma=GAMMA, axial=True)
    n = ctx.n
    ctx.set_state(np.ones(n), np.zeros(n), np.zeros(n), np.ones(n), np.full(n, 2.5))
    ctx.step(30)
    rho, u, v, p, e = ctx.state()
    assert np.abs(v).max() <= 1e-13 and np.abs(u).max() <= 1e-13 and np.abs(p - 1).max() <= 1e-13
    assert ctx.params()["axial"] == 1
    with pytest.raises(zfv.ZfvError):
        ctx.set_flags(0, 0.1)
        ctx.refine(1)
    planar = zfv.Context()
    planar.mesh_rectangle(8, 8)
    with pytest.raises(zfv.ZfvError):
        planar.configure(acc=1, axial=True)
