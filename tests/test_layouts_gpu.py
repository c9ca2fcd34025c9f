"""The three step-kernel layouts (one thread per cell, four lanes per cell,
one thread per cell with the main slots' loads batched) and the launch
autotuner give bit-identical results, on uniform meshes and on adapted
meshes whose hanging faces use the second sub-face slots.

The quad layout combines per-face terms with DPP quad broadcasts in slot
order, the batched layout sums the main slots before the sub-face slots;
any lane-layout or ordering mistake shows up as a bit difference.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def adapted_context(zfv, layout, acc, flux, limiter, seed, bounds=(2, 1, 3, 3), block=256):
    """A 24x20 mesh refined twice with seeded flags (so sub-face slots exist),
    four-quadrant data, the given layout on every stage."""
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(24, 20, bounds=bounds)
    ctx.configure(acc=acc, flux=flux, limiter=limiter, cfl=0.5, gamma=GAMMA)
    for st in range(3):
        ctx.set_launch(st, layout, block)
    rng = np.random.default_rng(seed)
    for _ in range(2):
        cells, _, _ = ctx.mesh_download()
        ctx.set_state(*problems.four_quadrant(cells["cx"], cells["cy"], seed))
        ctx.set_flags_array(rng.choice([-1, 0, 1], size=ctx.n, p=[0.1, 0.6, 0.3]).astype(np.int32))
        ctx.refine(2, slope=False)
    cells, _, _ = ctx.mesh_download()
    ctx.set_state(*problems.four_quadrant(cells["cx"], cells["cy"], seed))
    return ctx


def assert_same(a, b, what):
    for k, (x, y) in enumerate(zip(a, b)):
        bad = np.flatnonzero(x != y)
        assert bad.size == 0, f"{what}: component {k} differs at {bad[:5]}: {x[bad[:3]]} vs {y[bad[:3]]}"


@pytest.mark.parametrize("acc,flux,limiter,block", [
    (1, 3, "MC", 256), (1, 2, "MC", 128), (1, 1, "MC", 512),
    (2, 3, "MC", 256), (2, 3, "minmod", 512), (2, 3, "vanLeer", 128), (2, 2, "superbee", 256), (2, 1, "koren", 512),
])
@pytest.mark.parametrize("layout", ["QUAD", "BATCH"])
def test_layout_equals_cell_on_adapted_mesh(zfv, layout, acc, flux, limiter, block):
    cell = adapted_context(zfv, zfv.CELL, acc, flux, limiter, seed=11)
    other = adapted_context(zfv, getattr(zfv, layout), acc, flux, limiter, seed=11, block=block)
    assert cell.n == other.n
    _, slots, _ = cell.mesh_download()
    n = cell.n
    assert np.any(slots["boundary"].reshape(n, 8)[:, 4:] >= 0), "mesh has no sub-face slots"
    for step in range(6):
        cell.step(1)
        other.step(1)
        assert cell.dt() == other.dt(), f"dt differs at step {step}"
        assert_same(cell.state(), other.state(), f"state after step {step + 1}")
        assert np.array_equal(cell.wait(), other.wait())
    if acc == 2:
        assert_same(cell.grad(), other.grad(), "gradients")


@pytest.mark.parametrize("layout", ["QUAD", "BATCH"])
def test_layout_equals_cell_set_flags(zfv, layout):
    cell = adapted_context(zfv, zfv.CELL, 2, 3, "MC", seed=5)
    other = adapted_context(zfv, getattr(zfv, layout), 2, 3, "MC", seed=5)
    for ctx in (cell, other):
        ctx.step(3)
        ctx.set_flags(1, 0.1)
    assert np.array_equal(cell.flags(), other.flags())
    assert_same(cell.grad(), other.grad(), "limited gradients of set_flags")


def test_autotune_keeps_state_and_results(zfv):
    ref = adapted_context(zfv, zfv.CELL, 2, 3, "MC", seed=7)
    tun = adapted_context(zfv, zfv.QUAD, 2, 3, "MC", seed=7)
    ref.step(2)
    tun.step(2)
    before = tun.state()
    ms = tun.autotune(reps=3)
    assert ms.shape == (3, 4, 3) and np.all(ms[1:, :3] > 0.0) and np.all(ms[0] == 0.0)
    assert np.all(ms[:, 3] == 0.0), "the tiled step does not apply to an adapted mesh"
    assert_same(before, tun.state(), "autotune changed the state")
    for st in (1, 2):
        layout, block = tun.launch(st)
        assert layout in (zfv.CELL, zfv.QUAD, zfv.BATCH) and block in (128, 256, 512)
        assert ms[st, layout, (128, 256, 512).index(block)] == ms[st, :3].min()
    ref.step(4)
    tun.step(4)
    assert_same(ref.state(), tun.state(), "after autotune")
