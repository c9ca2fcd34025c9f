"""zfv_conserved_totals_noalloc, the entry point that allocates nothing,
against zfv_conserved_totals through the C ABI: the same sums in the same
order, so the results must be bit-identical, on two calls in a row (the
second on a smaller mesh, whose partial sums overwrite only the head of the
shared scratch), with the state held in curve order on an adapted mesh, and
again on the first mesh after the smaller one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def adapted(zfv, nx, ny, seed, cycles):
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, ny)
    ctx.configure(acc=2, flux="hllc", limiter="minmod", cfl=0.5, gamma=GAMMA)
    cells, _, _ = ctx.mesh_download()
    ctx.set_state(*problems.circle_blast(cells["cx"], cells["cy"], seed))
    for _ in range(cycles):
        ctx.set_flags(1, 0.1)
        ctx.refine(3, slope=True)
        ctx.step(4)
    return ctx


def test_noalloc_totals_equal_totals(zfv):
    big = adapted(zfv, 160, 128, 3, 4)
    assert big.n > 160 * 128
    first = big.totals_noalloc()
    assert np.array_equal(first, big.totals())
    small = adapted(zfv, 24, 20, 4, 2)
    assert small.n < big.n
    got = small.totals_noalloc()
    assert np.array_equal(got, small.totals())
    assert not np.array_equal(got, first)
    small.step(3)   # state back in curve order
    assert np.array_equal(small.totals_noalloc(), small.totals())
    assert np.array_equal(big.totals_noalloc(), first)
    assert np.all(np.isfinite(first)) and first[0] > 0.0
