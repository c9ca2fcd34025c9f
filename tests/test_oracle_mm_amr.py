"""Known answers for the oracle's multi-material adaptation (MmFluid
set_flags, compute_grad and the split / merge distributor).

The reference's tests do not cover this path, so each part is pinned by
what the mathematics or the definition fixes:
* flags: +1 for a mixed cell and for a pure cell with a face neighbour of
  another material (or mixed), -1 otherwise, recomputed by brute force from
  the state and the mesh's adjacency;
* gradient: exact for a linear field away from the boundary (the MC limiter
  leaves a linear profile alone), zero in and next to mixed cells with
  composite Riemann problems on, the LSM gradient there without them;
* merge: conservative — refining every cell and coarsening every family
  back keeps mass, momentum, energy and each partial mass to round-off
  (the merge sums Q V, and the split of a linear field is conservative);
* split of a pure uniform state gives that state back (P and T re-closed).
"""
import numpy as np
import pytest

from test_oracle_mm import AIR, GAS, WATER

HEAVY = (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)


def _state(cx, cy, mats, disc=True):
    rho0, rho1 = mats[0][5], mats[1][5]
    a = ((cx - 0.5) ** 2 + (cy - 0.45) ** 2 < 0.2 ** 2).astype(float) if disc else np.zeros_like(cx)
    a = np.where((np.abs(np.hypot(cx - 0.5, cy - 0.45) - 0.2) < 0.03), 0.5, a) if disc else a
    rho = (1.0 - a) * rho0 + a * rho1
    b1 = a * rho1 / rho
    p = 1.0 + 0.1 * cx
    return rho, 0.1 + 0.05 * cy, -0.05 * cx, p, np.stack([1.0 - b1, b1])


def _setup(oracle, nx, mats, crp=True, disc=True):
    m = oracle.Mesh.rectangle(nx, nx)
    s = oracle.MmFluid(m, mats, 0.5, crp=crp)
    c = m.cells()
    s.set_state(*_state(c["cx"], c["cy"], mats, disc))
    return m, s


def _totals(m, s):
    z = s.state()
    vol = m.cells()["volume"]
    E = z["rho"] * (z["e"] + 0.5 * (z["u"] ** 2 + z["v"] ** 2))
    return np.array([np.sum(vol * z["rho"]), np.sum(vol * z["rho"] * z["u"]), np.sum(vol * z["rho"] * z["v"]),
                     np.sum(vol * E), *(np.sum(vol * z["rho"] * z["beta"][i]) for i in range(2))])


def test_flags_by_brute_force(oracle):
    m, s = _setup(oracle, 24, [GAS, HEAVY])
    s.set_flags()
    flags = m.cells()["flag"]
    z = s.state()
    sl = m.slots()
    bnd, adj = sl["boundary"].reshape(-1, 8), sl["adj"].reshape(-1, 8)
    present = z["beta"] > 0.0
    mixed = present.sum(axis=0) != 1
    index = np.where(mixed, -1, np.argmax(present, axis=0))
    expect = np.full(m.n, -1)
    for c in range(m.n):
        if mixed[c]:
            expect[c] = 1
            continue
        for k in range(8):
            if bnd[c, k] in (0, 3) and index[adj[c, k]] != index[c]:   # inner or periodic face
                expect[c] = 1
    assert np.array_equal(flags, expect)
    assert np.any(flags == 1) and np.any(flags == -1)


def test_gradient_of_a_linear_field_and_the_zero_rule(oracle):
    # one material: u = 0.1 + 0.05 y, v = -0.05 x, rho and p linear in x
    m, s = _setup(oracle, 16, [GAS, HEAVY], disc=False)
    c = m.cells()
    rho = 1.0 + 0.25 * c["cx"] - 0.125 * c["cy"]
    s.set_state(rho, 0.1 + 0.05 * c["cy"], -0.05 * c["cx"], 1.0 + 0.1 * c["cx"], np.stack([np.ones(m.n), np.zeros(m.n)]))
    s.compute_grad()
    gx, gy = s.grad()
    inner = (c["cx"] > 1.5 / 16) & (c["cx"] < 1 - 1.5 / 16) & (c["cy"] > 1.5 / 16) & (c["cy"] < 1 - 1.5 / 16)
    np.testing.assert_allclose(gx[0][inner], 0.25, rtol=0, atol=1e-13)
    np.testing.assert_allclose(gy[0][inner], -0.125, rtol=0, atol=1e-13)
    np.testing.assert_allclose(gy[1][inner], 0.05, rtol=0, atol=1e-13)
    np.testing.assert_allclose(gx[2][inner], -0.05, rtol=0, atol=1e-13)
    assert np.all(gx[4:6] == 0.0) and np.all(gy[4:6] == 0.0)   # pure: beta is constant
    # rows 6, 7: P = 1 + 0.1 x, T (nonlinear); rows 8, 9: the true density of
    # material 0 equals rho, material 1 is absent (NaN)
    np.testing.assert_allclose(gx[6][inner], 0.1, rtol=0, atol=1e-13)
    np.testing.assert_allclose(gy[6][inner], 0.0, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(gx[8], gx[0])
    np.testing.assert_array_equal(gy[8], gy[0])
    assert np.all(np.isnan(gx[9])) and np.all(np.isnan(gy[9]))
    # two materials: zero in and next to mixed cells only with CRP on
    for crp in (True, False):
        m2, s2 = _setup(oracle, 24, [GAS, HEAVY], crp=crp)
        s2.compute_grad()
        gx2, _ = s2.grad()
        z = s2.state()
        mixed = (z["beta"] > 0.0).sum(axis=0) > 1
        sl = m2.slots()
        bnd, adj = sl["boundary"].reshape(-1, 8), sl["adj"].reshape(-1, 8)
        near = mixed.copy()
        for cc in range(m2.n):
            for k in range(8):
                if bnd[cc, k] in (0, 3) and mixed[adj[cc, k]]:
                    near[cc] = True
        if crp:
            assert np.all(gx2[:, near] == 0.0)
            assert np.any(gx2[:, ~near] != 0.0)
        else:
            assert np.any(gx2[:, mixed] != 0.0)


@pytest.mark.parametrize("mats", [[GAS, HEAVY], [WATER, AIR]])
def test_refine_all_then_coarsen_all_conserves(oracle, mats):
    m, s = _setup(oracle, 16, mats)
    if mats[0] is WATER:
        c = m.cells()
        r, u, v, p, b = _state(c["cx"], c["cy"], mats)
        s.set_state(r, 30 * u, 30 * v, 1e5 * p, b)
    t0 = _totals(m, s)
    n0 = m.n
    s.compute_grad()
    m.set_flags(np.ones(m.n, np.int32))
    assert s.refine(1) == 1 and m.n == 4 * n0
    t1 = _totals(m, s)
    # the split is conservative for mass and momentum of a linear profile
    # (limited gradients, children at +-h/4): to round-off
    np.testing.assert_allclose(t1[:3], t0[:3], rtol=1e-13, atol=1e-13 * abs(t0[0]))
    m.set_flags(-np.ones(m.n, np.int32))
    assert s.refine(1) == 1 and m.n == n0
    t2 = _totals(m, s)
    # the merge is conservative for every total of the children
    np.testing.assert_allclose(t2, t1, rtol=1e-13, atol=1e-13 * abs(t1[0]))


def test_split_of_a_uniform_pure_state(oracle):
    m = oracle.Mesh.rectangle(8, 8)
    s = oracle.MmFluid(m, [WATER, AIR], 0.5)
    n = m.n
    s.set_state(np.full(n, 1000.0), np.full(n, 2.0), np.full(n, -1.0), np.full(n, 1.0e5),
                np.stack([np.ones(n), np.zeros(n)]))
    z0 = s.state()
    s.compute_grad()
    m.set_flags(np.ones(n, np.int32))
    assert s.refine(1) == 1
    z1 = s.state()
    for k in ("rho", "u", "v", "e"):
        assert np.all(z1[k] == z0[k][0]), k
    np.testing.assert_allclose(z1["P"], 1.0e5, rtol=1e-12)
    np.testing.assert_allclose(z1["T"], z0["T"][0], rtol=1e-13)
    assert np.all(z1["rhos"] == 0.0)   # the default-constructed ScalarSet of a split child
