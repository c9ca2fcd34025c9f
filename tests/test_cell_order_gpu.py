"""Curve order of the step kernels on refined meshes (zfv_set_cell_order).

On a mesh with refined cells the steps run on a copy of the mesh sorted
along the base-cell / Morton curve, and the state moves between the two
orders at the entry points.  A cell's update and the CFL minimum do not
depend on where the cell is stored, so every observable (state, dt, wait
counters, gradients, flags, mesh, conserved totals) must be bit-identical to
a context that keeps the mesh's own order throughout, through adaptations,
graph replays, the autotuner and both schemes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def refined_context(zfv, curve, acc=2, limiter="minmod", nx=48, ny=40, bounds=(2, 2, 2, 2), seed=3):
    """Circle data on an nx x ny base, adapted three times to max level 3."""
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, ny, bounds=bounds)
    ctx.configure(acc=acc, flux="hllc", limiter=limiter, cfl=0.5, gamma=GAMMA)
    ctx.set_cell_order(curve)
    for _ in range(4):
        cells, _, _ = ctx.mesh_download()
        ctx.set_state(*problems.circle_blast(cells["cx"], cells["cy"], seed))
        ctx.set_flags(1, 0.1)
        ctx.refine(3, slope=False)
    cells, _, _ = ctx.mesh_download()
    ctx.set_state(*problems.circle_blast(cells["cx"], cells["cy"], seed))
    return ctx


def assert_same(a, b, what):
    sa, sb = a.state(), b.state()
    for k, (x, y) in enumerate(zip(sa, sb)):
        bad = np.flatnonzero(x != y)
        assert bad.size == 0, f"{what}: component {k} differs at {bad[:5]}: {x[bad[:3]]} vs {y[bad[:3]]}"
    assert a.dt() == b.dt(), f"{what}: dt"
    assert np.array_equal(a.wait(), b.wait()), f"{what}: wait"


def adapt(ctx):
    ctx.set_flags(1, 0.1)
    ctx.refine(3, slope=True)


@pytest.mark.parametrize("acc", [1, 2])
def test_curve_order_equals_mesh_order_through_adaptations(zfv, acc):
    cur = refined_context(zfv, True, acc=acc)
    ref = refined_context(zfv, False, acc=acc)
    assert cur.n == ref.n and cur.n > 48 * 40
    for cycle in range(6):
        cur.step(4)
        ref.step(4)
        assert cur.cell_order() == (1, 1), "steps on a refined mesh run in curve order"
        assert ref.cell_order() == (0, 0)
        if acc == 2 and cycle == 2:
            gc, gr = cur.grad(), ref.grad()
            for x, y in zip(gc, gr):
                assert np.array_equal(x, y), "gradients of the last step"
        assert_same(cur, ref, f"cycle {cycle}")
        assert np.array_equal(cur.totals(), ref.totals())
        adapt(cur)
        adapt(ref)
        assert np.array_equal(cur.flags(), ref.flags()), f"cycle {cycle}: flags"
        cc, cs, _ = cur.mesh_download()
        rc, rs, _ = ref.mesh_download()
        for k in ("cx", "cy", "level", "z_idx"):
            assert np.array_equal(cc[k], rc[k]), f"cycle {cycle}: mesh {k}"
        assert np.array_equal(cs["adj"], rs["adj"]), f"cycle {cycle}: adjacency"
        assert_same(cur, ref, f"after adaptation {cycle}")


def test_slope_distributor_reads_gradients_of_curve_steps(zfv):
    """refine(slope) right after steps (no set_flags in between) reads the
    last step's gradients, which were written in curve order."""
    cur = refined_context(zfv, True)
    ref = refined_context(zfv, False)
    for cycle in range(3):
        cur.step(3)
        ref.step(3)
        for c in (cur, ref):
            c.set_flags_array(np.where(np.arange(c.n) % 7 == 0, 1, 0).astype(np.int32))
            c.refine(3, slope=True)
        assert_same(cur, ref, f"cycle {cycle}")


def test_curve_order_with_graphs_and_autotune(zfv):
    cur = refined_context(zfv, True)
    ref = refined_context(zfv, False)
    cur.graph_enable(True)
    cur.step(2)
    ref.step(2)
    ms = cur.autotune(reps=2)
    assert ms.shape == (3, 4, 3)
    for k in (13, 6, 7):
        cur.step(k)
        ref.step(k)
        assert_same(cur, ref, f"after {k} steps with graphs")
    adapt(cur)
    adapt(ref)
    cur.step(12)
    ref.step(12)
    assert_same(cur, ref, "graphs after an adaptation")


def test_state_and_wait_uploads_in_curve_mode(zfv):
    """Uploads between steps land in the mesh's order."""
    from zephyr_amd import problems
    cur = refined_context(zfv, True)
    ref = refined_context(zfv, False)
    cur.step(2)
    ref.step(2)
    cells, _, _ = ref.mesh_download()
    st = problems.circle_blast(cells["cx"], cells["cy"], 9)
    w = np.where(np.arange(ref.n) % 11 == 0, 2, 0).astype(np.int32)
    for c in (cur, ref):
        c.set_state(*st)
        c.set_wait(w)
        c.step(5)
    assert_same(cur, ref, "after uploads")


def test_mode_switch_and_rectangles(zfv):
    """Switching to the mesh's order mid-run changes nothing; a rectangle
    (no refined cells) never enters curve order."""
    cur = refined_context(zfv, True)
    ref = refined_context(zfv, False)
    cur.step(3)
    ref.step(3)
    cur.set_cell_order(False)
    assert cur.cell_order() == (0, 0)
    cur.step(3)
    ref.step(3)
    assert_same(cur, ref, "after switching off")
    cur.set_cell_order(True)
    cur.step(3)
    ref.step(3)
    assert cur.cell_order() == (1, 1)
    assert_same(cur, ref, "after switching on")
    rect = zfv.Context()
    rect.mesh_rectangle(32, 32)
    rect.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
    from zephyr_amd import problems
    cells, _, _ = rect.mesh_download()
    rect.set_state(*problems.four_quadrant(cells["cx"], cells["cy"], 1))
    rect.step(2)
    assert rect.cell_order() == (1, 0)


def curve_keys(cells, lmax=3):
    """Host restatement of the curve key: base cell, then the Morton index of
    the cell's corner at level lmax inside it."""
    L = cells["level"].astype(np.int64)
    return (cells["b_idx"].astype(np.int64) << (2 * lmax)) | (cells["z_idx"].astype(np.int64) << (2 * (lmax - L)))


@pytest.mark.parametrize("every", [1, 4])
def test_carried_curve_order_equals_rebuilt_order(zfv, every):
    """The order carried across each adaptation (children where their parent
    was, a parent where its children were) is the sorted curve order, and
    equals the order computed afresh after each adaptation."""
    carried = refined_context(zfv, True, nx=40, ny=36, seed=5)
    rebuilt = refined_context(zfv, True, nx=40, ny=36, seed=5)
    rebuilt.set_cell_order(True, rebuild=True)
    for cycle in range(6):
        for ctx in (carried, rebuilt):
            ctx.step(every)
            ctx.set_flags(1, 0.1)
            ctx.refine(3, slope=True)
        cells, _, _ = carried.mesh_download()
        pc, pr = carried.curve_order(), rebuilt.curve_order()
        assert np.array_equal(pc, pr), f"cycle {cycle}: carried order differs from the rebuilt one"
        assert np.array_equal(np.sort(pc), np.arange(carried.n)), f"cycle {cycle}: not a permutation"
        keys = curve_keys(cells)[pc]
        assert np.all(np.diff(keys) > 0), f"cycle {cycle}: not sorted along the curve"
        assert_same(carried, rebuilt, f"cycle {cycle}")
    assert carried.cell_order()[0] == 1 and rebuilt.cell_order()[0] == 2
