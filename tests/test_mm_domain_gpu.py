"""MmFluid::domain (mm_fluid.cpp:686-712) through zfv_mm_domain: after
interface_recovery, each cell holding material idx gives the whole cell or
the cell clipped by its PLIC line (Polygon::clip, polygon.cpp:266-292).

Fixed by the geometry, not by the formulas: the PLIC line of a mixed cell
cuts off exactly the volume fraction, so every clipped polygon's area is
alpha V to round-off; pure cells of the material are whole squares, cells
without it give nothing (and a mixed cell without an interface normal is
whole above alpha = 1/2, as the reference takes it); polygons are
counter-clockwise and lie in their cells; the domains of two materials
tile the square.  (A kept side of the wrong sign would give (1 - alpha) V.)
Also on an adapted mesh, after steps.
"""
import numpy as np
import pytest

from test_mm_gpu import GAS, HEAVY, WATER, AIR

pytestmark = pytest.mark.gpu


def shoelace(x, y):
    return 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)


def polygons(off, xs, ys):
    return [(xs[off[k]:off[k + 1]], ys[off[k]:off[k + 1]]) for k in range(len(off) - 1)]


def alpha_of(z, idx):
    with np.errstate(invalid="ignore", divide="ignore"):
        a = z["beta"][idx] * z["rho"] / z["rhos"][idx]
    return np.clip(np.nan_to_num(a, nan=0.0), 0.0, 1.0)


def check_domain(ctx, idx):
    off, xs, ys = ctx.mm_domain(idx)
    z = ctx.mm_state()
    nx, ny, _ = ctx.mm_interface()
    cells, _, (vx, vy) = ctx.mesh_download()
    vol = cells["volume"]
    a = alpha_of(z, idx)
    zero_n = (np.abs(nx[idx]) <= 1e-12) & (np.abs(ny[idx]) <= 1e-12)
    # cells that give a polygon, in cell order, and the area it must have:
    # the clip of a mixed cell with a normal is exact (alpha V); a cell
    # without a normal is taken whole above alpha = 1/2
    whole = (a >= 1.0) | ((a > 0.5) & zero_n)
    keep = ~((a <= 0.0) | ((a < 0.5) & zero_n))
    cells_k = np.flatnonzero(keep)
    want = np.where(whole, vol, a * vol)[cells_k]
    polys = polygons(off, xs, ys)
    assert len(polys) == len(cells_k)
    areas = np.array([shoelace(x, y) for x, y in polys])
    assert np.all(areas > -1e-15), "counter-clockwise polygons"
    np.testing.assert_allclose(areas, want, rtol=0, atol=1e-12 * vol.max())
    vx, vy = vx.reshape(-1, 9), vy.reshape(-1, 9)
    for (x, y), c in zip(polys, cells_k):
        assert x.min() >= vx[c, 0] - 1e-12 and x.max() <= vx[c, 8] + 1e-12
        assert y.min() >= vy[c, 0] - 1e-12 and y.max() <= vy[c, 8] + 1e-12
    mixed = keep & ~whole
    assert mixed.any(), "the state has mixed cells"
    return areas, a, vol, np.flatnonzero(mixed & ~zero_n).size


def test_planar_front(zfv):
    n = 32
    ctx = zfv.Context()
    ctx.mesh_rectangle(n, n)
    ctx.mm_configure([GAS, HEAVY], 0.5)
    i, j = np.divmod(np.arange(n * n), n)
    cx = (i + 0.5) / n
    # material 0 left of x = 0.4 + 0.1 y (a slanted front), mixed cells in between
    x0 = 0.4 + 0.1 * (j + 0.5) / n
    b0 = np.clip((x0 - (cx - 0.5 / n)) * n, 0.0, 1.0)
    st = (np.ones(n * n), np.zeros(n * n), np.zeros(n * n), np.ones(n * n), np.stack([b0, 1.0 - b0]))
    ctx.mm_set_state(*st)
    a0, al0, vol, k0 = check_domain(ctx, 0)
    a1, al1, _, k1 = check_domain(ctx, 1)
    assert k0 == k1 == n, "one clipped cell per row"
    assert abs(a0.sum() + a1.sum() - 1.0) <= 1e-12, "the two domains tile the square"
    with pytest.raises(zfv.ZfvError):
        ctx.mm_domain(2)


def test_bubble_on_an_adapted_mesh_after_steps(zfv):
    from zephyr_amd import problems
    n = 64
    ctx = zfv.Context()
    ctx.mesh_rectangle(n, n)
    ctx.mm_configure([WATER, AIR], 0.5)
    i, j = np.divmod(np.arange(n * n), n)
    ctx.mm_set_state(*problems.mm_bubble_shock((i + 0.5) / n, (j + 0.5) / n, 1.0 / n, 3))
    for _ in range(2):
        ctx.mm_set_flags()
        ctx.mm_refine(2)
    ctx.mm_step(5)
    areas, a, vol, k = check_domain(ctx, 1)   # the air bubble
    assert k > 50
    r = np.sqrt(areas.sum() / np.pi)
    assert 0.12 < r < 0.18, r
