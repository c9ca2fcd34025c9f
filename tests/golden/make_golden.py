"""Writes tests/golden/smf_small.npz: small seeded SmFluid runs computed by the
oracle (the reference itself cannot be built here: Eigen3, Boost and
<format> are absent), kept as regression fixtures for both the oracle and
the HIP path.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O  # noqa: E402

CASES = [
    # name, nx, ny, bounds, acc, flux, limiter, steps, seed
    ("fo_hllc_zoe", 16, 12, (2, 2, 2, 2), 1, O.HLLC, "MC", 12, 101),
    ("fo_rusanov_wall", 12, 16, (1, 1, 1, 1), 1, O.RUSANOV, "MC", 12, 102),
    ("muscl_mc_hllc", 16, 12, (2, 2, 2, 2), 2, O.HLLC, "MC", 10, 103),
    ("muscl_minmod_hll_periodic", 14, 10, (3, 3, 3, 3), 2, O.HLL, "minmod", 10, 104),
]


def run_case(nx, ny, bounds, acc, flux, limiter, steps, seed):
    m = O.Mesh.rectangle(nx, ny, bounds=bounds)
    s = O.SmFluid(m, acc=acc, flux=flux, limiter=limiter)
    init = m.quadrant_state(seed)
    s.set_state(*init)
    dts = []
    for _ in range(steps):
        assert s.update(1) == 0
        dts.append(s.dt())
    return np.array(init), np.array(s.state()), np.array(dts)


def main():
    out = {}
    for name, nx, ny, bounds, acc, flux, lim, steps, seed in CASES:
        init, fin, dts = run_case(nx, ny, bounds, acc, flux, lim, steps, seed)
        out[name + "/params"] = np.array([nx, ny, *bounds, acc, flux, O.LIMITERS.index(lim), steps, seed])
        out[name + "/init"] = init
        out[name + "/final"] = fin
        out[name + "/dt"] = dts
    np.savez_compressed(os.path.join(HERE, "smf_small.npz"), **out)
    print("wrote", os.path.join(HERE, "smf_small.npz"))


if __name__ == "__main__":
    main()
