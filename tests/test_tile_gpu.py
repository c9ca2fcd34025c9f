"""The tiled MUSCL step (one kernel per step, half states and gradients in
LDS; layout TILE of stage 1) against the gradient + corrector kernels on
the same Rectangle meshes: states, time steps, isolation counters and the
last step's gradients bit-identical, for every tile shape, every domain
side (transmissive, wall, periodic pairs, a one-sided periodic side),
meshes smaller than a tile and meshes whose sides are not multiples of the
tile's, several limiters, batches of problems, graph replays, and the
isolation path.  The tiled step is the default on Rectangle meshes with
MUSCL + HLLC, so the oracle parity tests run it as well."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4
SHAPES = (512, 768, 1024)


def quadrants(nx, ny, seed):
    from zephyr_amd import problems
    i, j = np.divmod(np.arange(nx * ny), ny)
    return problems.four_quadrant((i + 0.5) / nx, (j + 0.5) / ny, seed)


def pair(zfv, nx, ny, bounds, limiter, shape, seed=3):
    out = []
    for tiled in (False, True):
        ctx = zfv.Context()
        ctx.mesh_rectangle(nx, ny, bounds=bounds)
        ctx.configure(acc=2, flux="hllc", limiter=limiter, cfl=0.5, gamma=GAMMA)
        if tiled:
            ctx.set_launch(1, zfv.TILE, shape)
            assert ctx.launch(1) == (zfv.TILE, shape)
        else:
            ctx.set_launch(1, zfv.CELL, 256)
        ctx.set_state(*quadrants(nx, ny, seed))
        out.append(ctx)
    return out


def assert_same(a, b, what):
    for k, (x, y) in enumerate(zip(a, b)):
        bad = np.flatnonzero(x != y)
        assert bad.size == 0, f"{what}: component {k} differs at {bad.size} cells, first {bad[:4]}"


def run_compare(ref, til, steps, what):
    for k in range(steps):
        ref.step(1)
        til.step(1)
        assert ref.dt() == til.dt(), f"{what}: dt at step {k}"
        assert_same(ref.state(), til.state(), f"{what}: state after step {k + 1}")
        assert np.array_equal(ref.wait(), til.wait()), f"{what}: wait after step {k + 1}"
    assert_same(ref.grad(), til.grad(), f"{what}: gradients of the last step")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("nx,ny,bounds", [
    (40, 70, (2, 2, 2, 2)), (13, 130, (1, 1, 1, 1)), (31, 9, (3, 3, 3, 3)),
    (20, 64, (3, 3, 2, 1)), (17, 62, (2, 1, 3, 3)), (9, 125, (3, 2, 1, 2)),
])
def test_tiled_step_equals_stage_kernels(zfv, shape, nx, ny, bounds):
    ref, til = pair(zfv, nx, ny, bounds, "MC", shape)
    run_compare(ref, til, 5, f"{nx}x{ny} {bounds}")


@pytest.mark.parametrize("nx,ny", [(1, 1), (1, 7), (5, 1), (2, 2), (3, 66), (7, 3)])
@pytest.mark.parametrize("bounds", [(2, 2, 2, 2), (3, 3, 3, 3)])
def test_tiled_step_on_tiny_meshes(zfv, nx, ny, bounds):
    ref, til = pair(zfv, nx, ny, bounds, "minmod", 512, seed=8)
    run_compare(ref, til, 3, f"{nx}x{ny} {bounds}")


@pytest.mark.parametrize("limiter", ["none", "unit", "minmod", "vanLeer", "vanAlbada1", "CHARM", "HCUS",
                                     "HQUICK", "koren", "smart", "superbee", "vanAlbada2"])
def test_tiled_step_limiters(zfv, limiter):
    ref, til = pair(zfv, 48, 80, (2, 1, 3, 3), limiter, 768, seed=5)
    run_compare(ref, til, 4, limiter)


def test_tiled_step_is_the_default_and_autotuned(zfv):
    ctx = zfv.Context()
    ctx.mesh_rectangle(64, 96)
    ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
    assert ctx.launch(1)[0] == zfv.TILE
    ctx.set_state(*quadrants(64, 96, 2))
    ref = zfv.Context()
    ref.mesh_rectangle(64, 96)
    ref.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
    ref.set_launch(1, zfv.CELL, 256)
    ref.set_state(*quadrants(64, 96, 2))
    ctx.step(2)
    ref.step(2)
    before = ctx.state()
    ms = ctx.autotune(reps=2)
    assert np.all(ms[1, 3] > 0.0)
    assert_same(before, ctx.state(), "autotune changed the state")
    ctx.graph_enable(True)
    for k in (7, 6, 13):
        ctx.step(k)
        ref.step(k)
        assert_same(ref.state(), ctx.state(), f"after {k} more steps (graphs, autotuned)")
        assert ref.dt() == ctx.dt()
    # a scheme the tiled kernel does not cover runs the stage kernels
    for c in (ctx, ref):
        c.configure(acc=2, flux="hll", limiter="MC", cfl=0.5, gamma=GAMMA)
        c.step(3)
    assert_same(ref.state(), ctx.state(), "HLL")


def test_tiled_step_on_batches(zfv):
    sizes = [(64, 64), (17, 9), (100, 130), (5, 70), (62, 62), (1, 1), (33, 200)]
    bounds = [(2, 2, 2, 2), (1, 2, 3, 3), (3, 3, 3, 3), (1, 1, 1, 1), (3, 3, 2, 1), (2, 2, 2, 2), (2, 1, 3, 3)]
    out = []
    for layout, block in ((zfv.CELL, 256), (zfv.TILE, 1024)):
        ctx = zfv.Context()
        ctx.batch_rectangles(sizes, bounds=bounds)
        ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
        ctx.set_launch(1, layout, block)
        begin, count = ctx.batch_layout()
        full = [np.ones(ctx.n), np.zeros(ctx.n), np.zeros(ctx.n), np.ones(ctx.n), np.full(ctx.n, 2.5)]
        for p, (nx, ny) in enumerate(sizes):
            st = quadrants(nx, ny, 40 + p)
            for k in range(5):
                full[k][begin[p]:begin[p] + count[p]] = st[k]
        ctx.set_state(*full)
        ctx.step(7)
        out.append((ctx.state(), ctx.batch_dt(), ctx.wait(), ctx.grad()))
    assert np.array_equal(out[0][1], out[1][1])
    assert_same(out[0][0], out[1][0], "batch state")
    assert np.array_equal(out[0][2], out[1][2])
    assert_same(out[0][3], out[1][3], "batch gradients")


def test_tiled_step_isolation_path(zfv):
    """A state whose predictor or corrector goes bad isolates cells (wait
    counters, robust HLL faces); the tiled step takes the same branches."""
    nx, ny = 40, 36
    st = list(quadrants(nx, ny, 4))
    i, j = np.divmod(np.arange(nx * ny), ny)
    hot = (np.abs(i - 20) < 2) & (np.abs(j - 18) < 2)
    st[0] = np.where(hot, 1e-3, st[0])
    st[3] = np.where(hot, 50.0, st[3])
    st[1] = np.where(hot, 3.0, st[1])
    from zephyr_amd import problems
    st[4] = problems.energy_rP(st[0], st[3], GAMMA)
    ref, til = pair(zfv, nx, ny, (2, 2, 2, 2), "MC", 512)
    waits = np.zeros(nx * ny, np.int32)
    waits[(np.abs(i - 10) < 3) & (np.abs(j - 10) < 3)] = 2
    for c in (ref, til):
        c.set_state(*st)
        c.set_wait(waits)
    run_compare(ref, til, 6, "isolation")
    assert ref.wait().max() > 0 or waits.max() > 0


def test_tiled_layout_refused_where_it_does_not_apply(zfv):
    ctx = zfv.Context()
    ctx.mesh_rectangle(16, 16)
    ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
    with pytest.raises(zfv.ZfvError):
        ctx.set_launch(1, zfv.TILE, 640)
    with pytest.raises(zfv.ZfvError):
        ctx.set_launch(2, zfv.TILE, 512)
    from zephyr_amd import problems
    cells, _, _ = ctx.mesh_download()
    ctx.set_state(*problems.circle_blast(cells["cx"], cells["cy"], 2))
    ctx.set_flags(1, 0.1)
    ctx.refine(2, slope=True)
    with pytest.raises(zfv.ZfvError):
        ctx.set_launch(1, zfv.TILE, 512)
    ctx.step(2)   # adapted mesh: the stage kernels


@pytest.mark.parametrize("shape", SHAPES)
def test_persistent_schedule_bit_identical(zfv, shape):
    """The tiled step by persistent workgroups (each walking the tiles of its
    XCD's range) against one workgroup per tile: a mesh of many tiles, a
    mesh smaller than the resident grid, and a batch of problems."""
    for nx, ny, bounds in ((300, 260, (2, 2, 3, 3)), (9, 40, (1, 2, 2, 1))):
        _, b = pair(zfv, nx, ny, bounds, "MC", shape, seed=9)
        ref = pair(zfv, nx, ny, bounds, "MC", shape, seed=9)[1]   # tiled, one workgroup per tile
        b.set_tile_schedule(True)
        assert b.tile_schedule() and not ref.tile_schedule()
        run_compare(ref, b, 4, f"persistent {nx}x{ny} shape {shape}")
    sizes = [(64, 64), (17, 9), (100, 130), (5, 70), (1, 1)]
    out = []
    for persistent in (False, True):
        ctx = zfv.Context()
        ctx.batch_rectangles(sizes)
        ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
        ctx.set_launch(1, zfv.TILE, shape)
        ctx.set_tile_schedule(persistent)
        begin, count = ctx.batch_layout()
        full = [np.ones(ctx.n), np.zeros(ctx.n), np.zeros(ctx.n), np.ones(ctx.n), np.full(ctx.n, 2.5)]
        for p, (nx, ny) in enumerate(sizes):
            st = quadrants(nx, ny, 60 + p)
            for k in range(5):
                full[k][begin[p]:begin[p] + count[p]] = st[k]
        ctx.set_state(*full)
        ctx.step(5)
        out.append((ctx.state(), ctx.batch_dt(), ctx.wait()))
    assert_same(out[0][0], out[1][0], "batch state, persistent schedule")
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
