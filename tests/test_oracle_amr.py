"""Pins of the oracle's AMR restatement (no reference test asserts an AMR
result; tests/amr/random.cpp only prints).  Checked here:
  * structural invariants after every adaptation (amr_checks.check_mesh):
    back-pointing neighbours, matching face centres, 2:1 balance, boundary
    slots referencing themselves, total volume = domain area;
  * balanced flags (check_flags): wanted levels differ by <= 1, whole
    sibling families coarsen;
  * the cell-count identity n_short = n_retain + 4 n_refine + n_coarse / 4
    (statistics.h:77-80);
  * refine-then-coarsen restores the vertices bit for bit (centres and
    volumes to round-off: they are recomputed by other formulas);
  * the "const" distributor with a conservative merge keeps mass; the slope
    distributor reproduces linear fields in the children;
  * a random-flag stress as tests/amr/random.cpp:20-36 (p = 0.80, 0.18, 0.02).
"""
import numpy as np
import pytest

from amr_checks import check_balanced_flags, check_mesh


def mesh_dicts(m):
    return m.cells(), m.slots()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_flags_stress(oracle, seed):
    m = oracle.Mesh.rectangle(12, 10, bounds=(2, 1, 2, 1))
    rng = np.random.default_rng(seed)
    for it in range(40):
        c, s = mesh_dicts(m)
        f = rng.choice([-1, 0, 1], size=m.n, p=[0.80, 0.18, 0.02]).astype(np.int32)
        if it < 6:
            f = rng.choice([-1, 0, 1], size=m.n, p=[0.2, 0.3, 0.5]).astype(np.int32)
        m.set_flags(f)
        m.balance_flags(3)
        bal = m.cells()["flag"]
        check_balanced_flags(c, s, bal, 3)
        n_ret, n_ref, n_crs = (bal == 0).sum(), (bal > 0).sum(), (bal < 0).sum()
        m.set_flags(f)
        m.refine(3)
        if n_ref + n_crs:
            assert m.n == n_ret + 4 * n_ref + n_crs // 4
        c, s = mesh_dicts(m)
        check_mesh(c, s, domain_area=1.0, max_level=3)


def test_refine_then_coarsen_restores_mesh(oracle):
    m = oracle.Mesh.rectangle(6, 5, 0.0, 1.2, 0.0, 1.0)
    c0, s0 = mesh_dicts(m)
    vx0, vy0 = m.verts()
    m.set_flags(np.ones(m.n, np.int32))
    m.refine(2)
    assert m.n == 4 * 30
    check_mesh(*mesh_dicts(m), domain_area=1.2, max_level=2)
    m.set_flags(-np.ones(m.n, np.int32))
    m.refine(2)
    c1, s1 = mesh_dicts(m)
    assert m.n == 30
    # same cells, possibly permuted: match by b_idx
    # (the restored centre is Quad::centroid_oxy and the volume the cross
    # product, while the base cells used the vertex average and hx*hy: equal
    # to round-off; the vertices are rebuilt from the same corners: bit-equal)
    order = np.argsort(c1["b_idx"])
    for k in ("cx", "cy", "volume"):
        np.testing.assert_allclose(c1[k][order], c0[k], rtol=1e-15, atol=0)
    vx1, vy1 = m.verts()
    assert np.array_equal(vx1.reshape(-1, 9)[order], vx0.reshape(-1, 9))
    assert np.array_equal(vy1.reshape(-1, 9)[order], vy0.reshape(-1, 9))
    assert np.array_equal(c1["level"], c0["level"])


def test_const_distributor_conserves_mass(oracle):
    m = oracle.Mesh.rectangle(8, 8)
    s = oracle.SmFluid(m, acc=2)
    s.set_state(*m.quadrant_state(4))
    t0 = s.totals()
    m.set_flags(np.ones(m.n, np.int32))
    s.refine(1, slope=False)
    t1 = s.totals()
    np.testing.assert_allclose(t1, t0, rtol=1e-14)
    m.set_flags(-np.ones(m.n, np.int32))
    s.refine(1, slope=False)
    np.testing.assert_allclose(s.totals(), t0, rtol=1e-14)


def test_slope_distributor_linear_field(oracle):
    """Children of a linear density field get the exact linear values."""
    m = oracle.Mesh.rectangle(10, 10)
    s = oracle.SmFluid(m, acc=2, limiter="unit")
    c = m.cells()
    rho = 2.0 + 0.3 * c["cx"] - 0.2 * c["cy"]
    p = np.ones(m.n)
    z = np.zeros(m.n)
    s.set_state(rho, z, z, p, p / (0.4 * rho))
    s.compute_grad()
    flags = np.zeros(m.n, np.int32)
    ii, jj = np.divmod(np.arange(m.n), 10)
    flags[(ii == 4) & (jj == 5)] = 1
    m.set_flags(flags)
    s.refine(1, slope=True)
    c = m.cells()
    r = s.state()[0]
    fine = c["level"] == 1
    assert fine.sum() == 4
    np.testing.assert_allclose(r[fine], 2.0 + 0.3 * c["cx"][fine] - 0.2 * c["cy"][fine], rtol=1e-13)


def test_relative_jump_flags(oracle):
    """set_flags criterion (sm_fluid.cpp:517-643): +1 at a 10 % jump, -1 in smooth data."""
    m = oracle.Mesh.rectangle(10, 4)
    s = oracle.SmFluid(m, acc=2)
    c = m.cells()
    rho = np.where(c["cx"] < 0.5, 1.0, 1.1)
    p = np.ones(m.n)
    z = np.zeros(m.n)
    s.set_state(rho, z, z, p, p / (0.4 * rho))
    s.set_flags(0)
    f = m.cells()["flag"]
    near = np.abs(c["cx"] - 0.5) < 0.1
    assert np.all(f[near] == 1) and np.all(f[~near] == -1)
    rho = np.where(c["cx"] < 0.5, 1.0, 1.03)   # 3 %: between 0.4*5 % and 5 %
    s.set_state(rho, z, z, p, p / (0.4 * rho))
    s.set_flags(0)
    f = m.cells()["flag"]
    assert np.all(f[near] == 0) and np.all(f[~near] == -1)
