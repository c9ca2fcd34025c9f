"""Batches of independent problems in one context (zfv_batch_rectangles):
every problem, advanced with its own CFL time step by the one-launch-per-stage
kernels, is bit-identical to the same problem run alone; graph-captured steps
are bit-identical to eager ones, also with eager work, max_dt changes and
downloads between replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4
SIZES = [(17, 9), (40, 31), (8, 64), (33, 33), (5, 7), (64, 48)]
BOUNDS = [(2, 2, 2, 2), (1, 2, 3, 3), (3, 3, 1, 1), (2, 1, 2, 1), (1, 1, 1, 1), (3, 3, 3, 3)]


def quadrant(nx, ny, seed):
    from zephyr_amd import problems
    i, j = np.divmod(np.arange(nx * ny), ny)
    return problems.four_quadrant((i + 0.5) / nx, (j + 0.5) / ny, seed)


def batch_context(zfv, acc, limiter, sizes=SIZES, bounds=BOUNDS):
    ctx = zfv.Context()
    ctx.batch_rectangles(sizes, bounds=bounds)
    ctx.configure(acc=acc, flux="hllc", limiter=limiter, cfl=0.5, gamma=GAMMA)
    begin, count = ctx.batch_layout()
    full = [np.ones(ctx.n) for _ in range(5)]
    full[1][:] = 0.0
    full[2][:] = 0.0
    full[4][:] = 1.0 / (GAMMA - 1.0)
    for p, (nx, ny) in enumerate(sizes):
        st = quadrant(nx, ny, 100 + p)
        for k in range(5):
            full[k][begin[p]:begin[p] + count[p]] = st[k]
    ctx.set_state(*full)
    return ctx, begin, count


@pytest.mark.parametrize("acc,limiter", [(1, "MC"), (2, "MC"), (2, "minmod")])
def test_batch_equals_problems_alone(zfv, acc, limiter):
    ctx, begin, count = batch_context(zfv, acc, limiter)
    assert np.all(begin[:-1] % 512 == 0)
    assert list(count) == [a * b for a, b in SIZES]
    singles = []
    for p, ((nx, ny), b) in enumerate(zip(SIZES, BOUNDS)):
        c = zfv.Context()
        c.mesh_rectangle(nx, ny, bounds=b)
        c.configure(acc=acc, flux="hllc", limiter=limiter, cfl=0.5, gamma=GAMMA)
        c.set_state(*quadrant(nx, ny, 100 + p))
        singles.append(c)
    for step in range(8):
        ctx.step(1)
        dts = ctx.batch_dt()
        st = ctx.state()
        for p, c in enumerate(singles):
            c.step(1)
            assert dts[p] == c.dt(), f"problem {p} dt at step {step}"
            sl = slice(begin[p], begin[p] + count[p])
            for k, a in enumerate(c.state()):
                assert np.array_equal(st[k][sl], a), f"problem {p} component {k} step {step}"
    assert len(set(dts.tolist())) > 1, "per-problem time steps should differ"


def test_batch_autotune_and_layouts(zfv):
    ref, begin, count = batch_context(zfv, 2, "MC")
    tun, _, _ = batch_context(zfv, 2, "MC")
    tun.set_launch(1, zfv.QUAD, 128)
    tun.set_launch(2, zfv.BATCH, 512)
    ref.step(3)
    tun.step(3)
    tun.set_launch(1, zfv.BATCH, 256)
    tun.set_launch(2, zfv.QUAD, 512)
    ref.step(2)
    tun.step(2)
    tun.autotune(reps=2)
    ref.step(5)
    tun.step(5)
    for a, b in zip(ref.state(), tun.state()):
        assert np.array_equal(a, b)
    assert np.array_equal(ref.batch_dt(), tun.batch_dt())


@pytest.mark.parametrize("acc", [1, 2])
def test_graph_replays_equal_eager(zfv, acc):
    eager, _, _ = batch_context(zfv, acc, "MC")
    graph, _, _ = batch_context(zfv, acc, "MC")
    graph.graph_enable(True)
    plan = [7, 6, 13, 1, 12, 6]   # replays of two phases, with eager remainders between
    for i, k in enumerate(plan):
        if i == 3:
            for c in (eager, graph):
                c.set_max_dt(1e-4)
        eager.step(k)
        graph.step(k)
        a, b = eager.state(), graph.state()
        for x, y in zip(a, b):
            assert np.array_equal(x, y), f"after call {i}"
        assert np.array_equal(eager.batch_dt(), graph.batch_dt())
        assert np.array_equal(eager.wait(), graph.wait())
    assert np.all(graph.batch_dt() <= 1e-4)


def test_padding_cells_keep_uploaded_values(zfv):
    """The step kernels never write a batch's padding cells; downloads
    return the uploaded padding values after odd and even step counts (the
    two state layers alternate).  A context that filled and freed device
    memory first makes stale values likely in fresh allocations."""
    junk = zfv.Context()
    junk.batch_rectangles(SIZES, bounds=BOUNDS)
    junk.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
    junk.set_state(*[np.full(junk.n, 7.5) for _ in range(5)])
    del junk
    ctx, begin, count = batch_context(zfv, 2, "MC")
    pad = np.ones(ctx.n, bool)
    for b, c in zip(begin[:-1], count):
        pad[b:b + c] = False
    assert pad.any()
    up = ctx.state()
    for k in (1, 1, 3):
        ctx.step(k)
        for a, b in zip(ctx.state(), up):
            assert np.array_equal(a[pad], b[pad])


def test_graph_on_single_problem_with_adaptation(zfv):
    """Graphs are re-captured after the mesh changes."""
    from zephyr_amd import problems
    out = []
    for use in (False, True):
        ctx = zfv.Context()
        ctx.mesh_rectangle(32, 32)
        ctx.configure(acc=2, flux="hllc", limiter="minmod", cfl=0.5, gamma=GAMMA)
        ctx.graph_enable(use)
        cells, _, _ = ctx.mesh_download()
        ctx.set_state(*problems.circle_blast(cells["cx"], cells["cy"], 3))
        for _ in range(3):
            ctx.set_flags(1, 0.1)
            ctx.refine(2, slope=True)
            ctx.step(6)
        out.append((ctx.state(), ctx.dt()))
    for x, y in zip(out[0][0], out[1][0]):
        assert np.array_equal(x, y)
    assert out[0][1] == out[1][1]
