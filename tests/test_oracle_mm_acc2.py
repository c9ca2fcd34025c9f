"""Known answers for the oracle's second-order multi-material update
(MmFluid::integrate with acc 2: compute_grad, fluxes_stage1 / stage2 and
PState::interpolation_update, mm_fluid.cpp:153-168, 521-680,
models.cpp:381-391).

The reference's tests do not cover this path; it is pinned by:
* one material: the second-order SmFluid step (sm_fluid.cpp, restated and
  pinned separately) bit for bit with the same dt — the Hancock predictor,
  the extrapolation of the half states with the gradients of the old layer,
  the boundary values and the conservative corrector all have to agree;
* a wall box: mass, energy and (without composite fluxes) every partial
  mass are conserved, which needs the face states seen from the two sides
  of each face to be the same extrapolations;
* every cell mixed with composite fluxes on: all gradients vanish, stage 1
  leaves mixed cells alone, so the step is the first-order step up to the
  re-closure of the face states (interpolation_update);
* a uniform mixture stays uniform.
"""
import numpy as np
import pytest

from test_oracle_mm import GAS, WALL, ZOE, _mesh_state

HEAVY = (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)


@pytest.mark.parametrize("limiter", ["MC", "minmod"])
def test_one_material_equals_single_fluid_second_order(oracle, limiter):
    m, (r, u, v, p, _) = _mesh_state(oracle, 24, 18, (WALL, ZOE, ZOE, WALL), [GAS], 5, mixed=False)
    mm = oracle.MmFluid(m, [GAS], cfl=0.5, crp=True)
    mm.set_accuracy(2)
    mm.set_limiter(limiter)
    mm.set_state(r, u, v, p, np.ones((1, m.n)))
    sm = oracle.SmFluid(m, gamma=1.4, flux=oracle.HLLC, limiter=limiter, acc=2, cfl=0.5)
    z = mm.state()
    sm.set_state(z["rho"], z["u"], z["v"], z["P"], z["e"])
    for st in (mm, sm):
        st.set_max_dt(1e-3)
    for step in range(6):
        assert mm.update() == 0
        sm.update(nthreads=1)
        assert mm.dt() == sm.dt() == 1e-3
        zm, zs = mm.state(), sm.state()
        for k, a in zip(("rho", "u", "v", "P", "e"), zs):
            assert np.array_equal(zm[k], a), f"{k} at step {step}"
    assert np.all(sm.wait() == 0)


def _totals(m, mm):
    z = mm.state()
    vol = m.cells()["volume"]
    E = z["rho"] * (z["e"] + 0.5 * (z["u"] ** 2 + z["v"] ** 2))
    return np.array([np.sum(vol * z["rho"]), np.sum(vol * E),
                     *(np.sum(vol * z["rho"] * z["beta"][i]) for i in range(2))])


@pytest.mark.parametrize("crp", [False, True])
def test_wall_box_conserves(oracle, crp):
    m, (r, u, v, p, beta) = _mesh_state(oracle, 20, 16, (WALL,) * 4, [GAS, HEAVY], 7)
    mm = oracle.MmFluid(m, [GAS, HEAVY], cfl=0.5, crp=crp)
    mm.set_accuracy(2)
    mm.set_state(r, u, v, p, beta)
    t0 = _totals(m, mm)
    for _ in range(10):
        assert mm.update() == 0
    t1 = _totals(m, mm)
    k = 4 if not crp else 2
    assert np.all(np.abs(t1 - t0)[:k] <= 1e-12 * np.abs(t0)[:k]), (t1 - t0) / t0
    # the second-order step differs from the first-order one
    m1, _ = _mesh_state(oracle, 20, 16, (WALL,) * 4, [GAS, HEAVY], 7)
    o1 = oracle.MmFluid(m1, [GAS, HEAVY], cfl=0.5, crp=crp)
    o1.set_state(r, u, v, p, beta)
    for _ in range(10):
        assert o1.update() == 0
    assert not np.allclose(o1.state()["rho"], mm.state()["rho"], rtol=1e-6, atol=0)


def test_all_mixed_with_composite_fluxes_is_the_first_order_step(oracle):
    m, (r, u, v, p, _) = _mesh_state(oracle, 16, 12, (ZOE, WALL, ZOE, ZOE), [GAS, HEAVY], 3)
    cl = m.cells()
    b0 = 0.3 + 0.4 * cl["cx"]   # strictly between 0 and 1: every cell mixed
    beta = np.stack([b0, 1.0 - b0])
    res = {}
    for acc in (1, 2):
        mm = oracle.MmFluid(m, [GAS, HEAVY], cfl=0.5, crp=True)
        mm.set_accuracy(acc)
        mm.set_state(r, u, v, p, beta)
        for _ in range(3):
            assert mm.update() == 0
        res[acc] = mm.state()
        if acc == 2:
            mm.compute_grad()
            gx, gy = mm.grad()
            assert np.all(gx == 0.0) and np.all(gy == 0.0)
    for k in ("rho", "u", "v", "P", "e"):
        np.testing.assert_allclose(res[2][k], res[1][k], rtol=1e-9, atol=1e-12, err_msg=k)


def test_uniform_mixture_stays_uniform(oracle):
    m = oracle.Mesh.rectangle(12, 10, bounds=(ZOE,) * 4)
    n = m.n
    for crp in (False, True):
        mm = oracle.MmFluid(m, [GAS, HEAVY], cfl=0.5, crp=crp)
        mm.set_accuracy(2)
        mm.set_state(np.full(n, 2.0), np.full(n, 0.3), np.full(n, -0.1), np.full(n, 1.5),
                     np.stack([np.full(n, 0.3), np.full(n, 0.7)]))
        z0 = mm.state()
        for _ in range(5):
            assert mm.update() == 0
        z1 = mm.state()
        for k in ("rho", "u", "v", "P"):
            assert np.allclose(z1[k], z0[k], rtol=1e-12, atol=1e-14), k
