"""The C++ problem file written against the host facade
(zephyr_amd/host/examples/smf_quadrants.cpp) gives the same result as the
same sequence of C ABI calls made from Python: Rectangle + SmFluid setters,
initial adaptation, then update / set_flags / refine per step."""
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def python_run(zfv, nx, steps, max_level, seed):
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, nx)
    ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=1.4)

    def init():
        cells, _, _ = ctx.mesh_download()
        ctx.set_state(*problems.four_quadrant(cells["cx"], cells["cy"], seed))

    for _ in range(max_level + 1):
        init()
        ctx.set_flags(0)
        ctx.refine(max_level, slope=True)
    init()
    time = 0.0
    for _ in range(steps):
        ctx.step(1)
        time += ctx.dt()
        if max_level > 0:
            ctx.set_flags(0)
            ctx.refine(max_level, slope=True)
    return ctx.n, time, ctx.dt(), ctx.totals()


def test_example_matches_python_path(zfv):
    from zephyr_amd import build as B
    nx, steps, max_level, seed = 40, 12, 2, 2024
    r = subprocess.run([B.EXAMPLE_PATHS["smf_quadrants"], str(nx), str(steps), str(max_level), str(seed)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    tok = r.stdout.split()
    cells, time, dt = int(tok[1]), float(tok[3]), float(tok[5])
    totals = np.array([float(t) for t in tok[7:11]])
    n, ptime, pdt, ptot = python_run(zfv, nx, steps, max_level, seed)
    assert cells == n and time == ptime and dt == pdt
    assert np.array_equal(totals, ptot)
