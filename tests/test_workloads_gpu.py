"""The benchmark workloads themselves against the oracle, through the C ABI.

* amr256_muscl (256 x 256 base on [0, 1]^2, up to 3 levels, MUSCL-Hancock +
  minmod + HLLC, circle blast of seed 2024, criterion 1 at 0.1, slope
  distributor, adaptation every 4 steps): set up and run exactly as
  bench.py runs it, for its 500 steps.  At every adaptation the flags, the
  cell order, levels, z / b indices, adjacency and every state component
  must be bit-identical to the oracle's, and after every step the dt.  The
  base spacing is a power of two, so every centre distance is exact and the
  LSM weights area / |drn|^3 are the same whether the cube is glibc's pow
  (the oracle, as gradient.h:256) or the correctly rounded cube the device
  uses after an adaptation: nothing leaves room for a last-bit difference.
* ensemble256_muscl (256 problems of s x s cells, s ~ U{64..512} from seed
  2024, per-problem four-quadrant data, MUSCL + MC + HLLC, per-problem dt):
  the whole batch, with the launch configuration autotuned and steps
  replayed from HIP graphs as bench.py runs it, against the oracle run on
  every problem alone: every problem's state and dt bit-identical.
* adjoint1024_muscl at its configured size (1024 x 1024, 200 steps,
  checkpoints every 20): the adjoint gradient against central differences
  of the device forward map on 64 seeded directions.
"""
import os

import numpy as np
import pytest

from amr_checks import check_mesh

pytestmark = pytest.mark.gpu

GAMMA = 1.4
SEED = 2024


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def bench_mod():
    import bench
    return bench


def test_amr256_workload_matches_oracle(oracle, zfv):
    B = bench_mod()
    name = "amr256_muscl"
    w = B.WORKLOADS[name]
    nt = threads()
    m = oracle.Mesh.rectangle(w["nx"], w["ny"])
    s = oracle.SmFluid(m, acc=w["acc"], flux=oracle.HLLC, limiter=w["limiter"], cfl=0.5)
    ctx = zfv.Context()
    ctx.mesh_rectangle(w["nx"], w["ny"])
    ctx.configure(acc=w["acc"], flux="hllc", limiter=w["limiter"], cfl=0.5, gamma=GAMMA)

    def o_centres():
        c = m.cells()
        return c["cx"], c["cy"]

    def d_centres():
        c, _, _ = ctx.mesh_download()
        return c["cx"], c["cy"]

    B.adapt_to_initial_data(s, name, SEED, o_centres, s.set_state,
                            lambda crit, thr: s.set_flags(crit, thr, nt),
                            lambda ml, slope: s.refine(ml, slope), w["max_level"], w["threshold"])
    B.adapt_to_initial_data(ctx, name, SEED, d_centres, ctx.set_state, ctx.set_flags,
                            lambda ml, slope: ctx.refine(ml, slope), w["max_level"], w["threshold"])

    def compare(what):
        cells, slots, _ = ctx.mesh_download()
        oc, osl = m.cells(), m.slots()
        assert len(cells["cx"]) == m.n, f"{what}: {len(cells['cx'])} cells vs {m.n}"
        for k in ("level", "b_idx", "z_idx"):
            assert np.array_equal(cells[k], oc[k]), f"{what}: {k}"
        assert np.array_equal(slots["boundary"], osl["boundary"]), f"{what}: boundary ids"
        defined = osl["boundary"] >= 0
        assert np.array_equal(slots["adj"][defined], osl["adj"][defined]), f"{what}: adjacency"
        for k, (a, b) in enumerate(zip(ctx.state(), s.state())):
            bad = np.flatnonzero(a != b)
            assert bad.size == 0, f"{what}: state component {k} differs at {bad.size} cells, first {bad[:4]}"
        return cells, slots

    compare("initial adaptation")
    levels = set()
    for g in range(w.get("steps", 500)):
        if g % w["every"] == 0:
            s.set_flags(1, w["threshold"], nt)
            ctx.set_flags(1, w["threshold"])
            assert np.array_equal(ctx.flags(), m.cells()["flag"]), f"step {g}: flags"
            s.refine(w["max_level"], True)
            ctx.refine(w["max_level"], slope=True)
            cells, _ = compare(f"adaptation before step {g}")
            levels.update(np.unique(cells["level"]).tolist())
        assert s.update(nt) == 0
        ctx.step(1)
        assert ctx.dt() == s.dt(), f"step {g}: dt {ctx.dt()!r} vs {s.dt()!r}"
    cells, slots = compare("after 500 steps")
    assert levels == {0, 1, 2, 3}
    # the same bits summed in a different order (the momentum totals are
    # round-off around zero: compared on the scale of the mass and energy)
    t_dev, t_orc = ctx.totals(), s.totals()
    np.testing.assert_allclose(t_dev, t_orc, rtol=1e-13, atol=1e-14 * np.abs(t_orc).max())
    check_mesh(cells, slots, domain_area=1.0, max_level=w["max_level"])


def test_ensemble256_workload_matches_oracle(oracle, zfv):
    B = bench_mod()
    w = B.WORKLOADS["ensemble256_muscl"]
    sizes = B.ensemble_sizes(w, SEED)
    assert len(sizes) == 256 and min(sizes) >= 64 and max(sizes) <= 512
    nt = threads()
    b = zfv.Context()
    b.batch_rectangles([(q, q) for q in sizes])
    b.configure(acc=w["acc"], flux="hllc", limiter=w["limiter"], cfl=0.5, gamma=GAMMA)
    begin, count = b.batch_layout()
    assert list(count) == [q * q for q in sizes]
    full = [np.ones(b.n), np.zeros(b.n), np.zeros(b.n), np.ones(b.n), np.full(b.n, 2.5)]
    for p, q in enumerate(sizes):
        st = B.ensemble_state(q, SEED, p)
        for k in range(5):
            full[k][begin[p]:begin[p] + count[p]] = st[k]
    b.set_state(*full)
    del full
    b.autotune(reps=1)       # the state does not advance
    b.graph_enable(True)
    steps = 8                # one 6-step graph replay and two eager steps
    b.step(steps)
    dts = b.batch_dt()
    got = b.state()
    b.close()
    assert len(set(dts.tolist())) > 100, "per-problem time steps should differ"
    for p, q in enumerate(sizes):
        m = oracle.Mesh.rectangle(q, q)
        f = oracle.SmFluid(m, acc=w["acc"], flux=oracle.HLLC, limiter=w["limiter"], cfl=0.5)
        f.set_state(*B.ensemble_state(q, SEED, p))
        for _ in range(steps):
            assert f.update(nt) == 0
        assert dts[p] == f.dt(), f"problem {p} ({q} x {q}): dt"
        sl = slice(begin[p], begin[p] + count[p])
        for k, a in enumerate(f.state()):
            bad = np.flatnonzero(got[k][sl] != a)
            assert bad.size == 0, f"problem {p} ({q} x {q}): component {k} differs at {bad.size} cells"
        del f, m


def test_adjoint1024_workload_matches_central_differences(zfv):
    """The adjoint workload at its configured size: 1024 x 1024, 200 steps,
    checkpoints every 20, J = integral of p^2 over [0.4, 0.6]^2; dJ/drho0
    (p, u, v held, e = e(rho, p)) on 64 seeded directions against central
    differences at eps 1e-7 (fewer limiter and wave-speed switches crossed
    than at 1e-6, rounding about 1e-16 J / eps).  Bar: the error over the
    RMS directional derivative at most 1e-5 on every direction; at the
    configured eps 1e-6 (truncation error larger, and a direction may cross
    a switch) a median of at most 1e-5 and 60 of the 64 directions within
    1e-4.  A dropped or wrong derivative term gives errors of 1e-3 and more
    on most directions."""
    B = bench_mod()
    from zephyr_amd import problems
    w = B.WORKLOADS["adjoint1024_muscl"]
    nx, ny, steps, every = w["nx"], w["ny"], w["steps"], w["every"]
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, ny)
    ctx.configure(acc=2, flux="hllc", limiter=w["limiter"], cfl=0.5, gamma=GAMMA)
    rho0, u0, v0, p0, _ = problems.smooth_flow_field(nx, ny, SEED)

    def state(rho):
        return rho, u0, v0, p0, problems.energy_rP(rho, p0, GAMMA)

    ctx.set_state(*state(rho0))
    J = ctx.adjoint_forward(steps, every, w["window"])
    lr, _, _, _, le = ctx.adjoint_backward()
    grad = lr - le * problems.energy_rP(rho0, p0, GAMMA) / rho0
    assert np.isfinite(grad).all() and J > 0.0
    ads, fds = [], {1e-6: [], 1e-7: []}
    for k in range(w["fd_dirs"]):
        d = problems.gaussian_random_field(nx, ny, 10_000 + k, corr_len=0.05)
        ads.append(float(np.dot(grad, d)))
        for eps in fds:
            ctx.set_state(*state(rho0 + eps * d))
            Jp = ctx.adjoint_forward(steps, steps, w["window"])
            ctx.set_state(*state(rho0 - eps * d))
            Jm = ctx.adjoint_forward(steps, steps, w["window"])
            fds[eps].append((Jp - Jm) / (2 * eps))
    ads = np.array(ads)
    rms = float(np.sqrt(np.mean(ads ** 2)))
    assert rms > 0.0
    for eps in fds:
        err = np.abs(np.array(fds[eps]) - ads) / rms
        print(f"eps {eps:g}: FD error / RMS derivative max {err.max():.3e} median {np.median(err):.3e}", flush=True)
    err7 = np.abs(np.array(fds[1e-7]) - ads) / rms
    assert err7.max() <= 1e-5, f"eps 1e-7: {err7.max():.3e}"
    # at eps 1e-6 a perturbation can cross a limiter or wave-speed switch
    # inside +-eps on a few directions; the bulk must still agree
    err6 = np.abs(np.array(fds[1e-6]) - ads) / rms
    assert np.median(err6) <= 1e-5 and np.sum(err6 <= 1e-4) >= 60, f"eps 1e-6: {np.sort(err6)[-4:]}"
