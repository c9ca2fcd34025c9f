"""The oracle's axisymmetric SmFluid (Rectangle::set_axial,
rectangle.cpp:97-103 and 590-602; SmFluid with m_axial,
sm_fluid.cpp:160-169, 236-245, 341-350), pinned by what the mathematics
fixes rather than by the formulas themselves:

* geometry: volume_alt is the y-moment of the cell, -sum (x2 - x1)
  (y1^2 + y1 y2 + y2^2) / 6 over the counter-clockwise boundary (Green's
  theorem), area_alt the y-moment of each face, and the shifted centres the
  y-weighted centroids int y^2 / int y; a bottom on the axis becomes a wall;
* a fluid at rest stays at rest (the pressure source balances the pressure
  force on the faces of different radius) at both orders and with every
  flux, to round-off; a uniform flow along the axis stays uniform;
* mass, x-momentum and energy weighted by volume_alt are conserved (the
  face areas telescope between neighbours; only the y-momentum has a
  source);
* the planar step on the same mesh (set_axial off) evolves a general state
  differently.
Each planted mistake -- no source term, the source without the predictor's
1/2, the planar volume in the update -- fails the resting-fluid check.
"""
import numpy as np
import pytest

WALL, ZOE, PERIODIC = 1, 2, 3


def axial_mesh(oracle, nx=12, ny=10, ymin=0.0, ymax=1.0, bounds=(ZOE, ZOE, ZOE, WALL)):
    return oracle.Mesh.rectangle(nx, ny, 0.0, 2.0, ymin, ymax, bounds=bounds, axial=True)


def test_geometry_is_the_y_moment(oracle):
    nx, ny, ymin, ymax = 7, 5, 0.25, 1.5
    m = axial_mesh(oracle, nx, ny, ymin, ymax)
    cells, slots, (vx, vy) = m.cells(), m.slots(), m.verts()
    vol_alt, area_alt = m.axial_geometry()
    vx, vy = vx.reshape(-1, 9), vy.reshape(-1, 9)
    ring = (0, 2, 8, 6)   # counter-clockwise corners of the SqQuad
    for c in range(m.n):
        V = 0.0
        for k in range(4):
            x1, y1 = vx[c, ring[k]], vy[c, ring[k]]
            x2, y2 = vx[c, ring[(k + 1) % 4]], vy[c, ring[(k + 1) % 4]]
            V -= (x2 - x1) * (y1 * y1 + y1 * y2 + y2 * y2) / 6.0
        assert abs(vol_alt[c] - V) <= 1e-14 * V
        y0, y1 = vy[c, 0], vy[c, 8]
        hx, hy = vx[c, 2] - vx[c, 0], y1 - y0
        # int y^2 dA / int y dA over the rectangle
        cyw = (y1 ** 3 - y0 ** 3) / 3.0 / ((y1 ** 2 - y0 ** 2) / 2.0)
        assert abs(cells["cy"][c] - cyw) <= 1e-14 * cyw
        ar = area_alt.reshape(-1, 8)[c]
        assert abs(ar[0] - hy * 0.5 * (y0 + y1)) <= 1e-15 and abs(ar[1] - ar[0]) == 0.0
        assert abs(ar[2] - hx * y0) <= 1e-15 and abs(ar[3] - hx * y1) <= 1e-15
        fy = slots["fcy"].reshape(-1, 8)[c]
        assert abs(fy[0] - cyw) <= 1e-14 and fy[2] == y0 and fy[3] == y1
    on_axis = axial_mesh(oracle, 4, 4, 0.0, 1.0, bounds=(ZOE, ZOE, ZOE, ZOE))
    b = on_axis.slots()["boundary"].reshape(-1, 8)
    assert np.all(b[np.arange(4) * 4, 2] == WALL), "a bottom on the axis is a wall"


def rest(oracle, acc, flux, limiter="MC", u=0.0, nx=12, ny=10, steps=20, plant=0, axial=True):
    m = axial_mesh(oracle, nx, ny, bounds=(PERIODIC, PERIODIC, WALL, WALL))
    s = oracle.SmFluid(m, acc=acc, flux=flux, limiter=limiter, cfl=0.4)
    s.set_axial(axial)
    n = m.n
    s.set_state(np.ones(n), np.full(n, u), np.zeros(n), np.ones(n), np.full(n, 1.0 / (0.4 * 1.0)))
    oracle.lib().orc_smf_plant_axial(plant)
    try:
        for _ in range(steps):
            assert s.update(2) == 0
    finally:
        oracle.lib().orc_smf_plant_axial(0)
    return s


@pytest.mark.parametrize("acc", [1, 2])
@pytest.mark.parametrize("flux", [0, 1, 2, 3, 4, 6])
def test_rest_stays_at_rest(oracle, acc, flux):
    s = rest(oracle, acc, flux)
    rho, u, v, p, e = s.state()
    assert np.abs(v).max() <= 1e-13 and np.abs(u).max() <= 1e-13
    assert np.abs(p - 1.0).max() <= 1e-13 and np.abs(rho - 1.0).max() <= 1e-13


@pytest.mark.parametrize("acc", [1, 2])
def test_uniform_axial_flow_stays_uniform(oracle, acc):
    s = rest(oracle, acc, 3, u=0.6)
    rho, u, v, p, e = s.state()
    assert np.abs(v).max() <= 1e-13 and np.abs(u - 0.6).max() <= 1e-13
    assert np.abs(p - 1.0).max() <= 1e-13


@pytest.mark.parametrize("acc", [1, 2])
def test_conservation_with_the_axial_volume(oracle, acc):
    m = axial_mesh(oracle, 16, 12, 0.0, 1.0, bounds=(PERIODIC, PERIODIC, WALL, WALL))
    s = oracle.SmFluid(m, acc=acc, flux=3, limiter="MC", cfl=0.4)
    s.set_axial(True)
    st = list(m.quadrant_state(5))
    s.set_state(*st)
    t0 = s.totals()
    vol_alt, _ = m.axial_geometry()
    for _ in range(8):
        assert s.update(2) == 0
    t1 = s.totals()
    for k in (0, 1, 3):   # mass, x-momentum, energy
        assert abs(t1[k] - t0[k]) <= 1e-13 * max(abs(t0[k]), 1.0), k
    rho = s.state()[0]
    assert abs(np.sum(rho * vol_alt) - t1[0]) <= 1e-12 * t1[0]
    assert abs(t1[2] - t0[2]) > 1e-6, "the y-momentum has a source"


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_mistakes_fail_the_rest_check(oracle, plant):
    s = rest(oracle, 2, 3, plant=plant)
    v = s.state()[2]
    assert np.abs(v).max() > 1e-6, f"planted mistake {plant} went unnoticed"


def test_planar_step_differs(oracle):
    """set_axial off on the axial mesh is the planar step on the shifted
    centres: a resting fluid stays at rest, and a general state evolves
    differently; set_axial needs the axial geometry; axial meshes are not
    adapted here."""
    s = rest(oracle, 1, 3, axial=False)
    assert np.abs(s.state()[2]).max() <= 1e-14
    out = []
    for ax in (False, True):
        m = axial_mesh(oracle, 12, 10, 0.0, 1.0, bounds=(PERIODIC, PERIODIC, WALL, WALL))
        s = oracle.SmFluid(m, acc=2, flux=3, cfl=0.4)
        s.set_axial(ax)
        s.set_state(*m.quadrant_state(2))
        for _ in range(3):
            assert s.update(2) == 0
        out.append(s.state()[2])
    assert np.abs(out[0] - out[1]).max() > 1e-3
    m = oracle.Mesh.rectangle(8, 8, bounds=(ZOE,) * 4)
    s = oracle.SmFluid(m)
    with pytest.raises(ValueError):
        s.set_axial(True)
    assert axial_mesh(oracle).refine(1) == -2, "axial meshes are not adapted"
