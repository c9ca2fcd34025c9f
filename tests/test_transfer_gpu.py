"""Device Transfer (zfv_transfer_*) against the oracle's restatement
(oracle/oracle_transfer.hpp, pinned in test_oracle_transfer.py) on the same
mesh, velocity samples and initial fraction.

Bar: bit-exact fraction, interface normals and (MUSCL) slopes after every
update, for every built method, with direction splitting and unsplit, on
periodic and open boxes; compute_dt and set_flags equal; the methods that
are not built are refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ZOE, PERIODIC = 2, 3
BUILT = ["CRP_V3", "CRP_V5", "CRP_SE", "CRP_N1", "MUSCLd", "MUSCLd_CRP", "MUSCL_MC", "MUSCL_MC_CRP", "WENO",
         "WENO_CRP"]


def rotation(x, y):
    return -(y - 0.45), (x - 0.5)


def pair(oracle, zfv, method, nx=48, ny=40, bounds=(PERIODIC,) * 4, field=rotation, cfl=0.5):
    m = oracle.Mesh.rectangle(nx, ny, bounds=bounds)
    cl, sl = m.cells(), m.slots()
    fvx, fvy = field(sl["fcx"], sl["fcy"])
    cvx, cvy = field(cl["cx"], cl["cy"])
    T = oracle.Transfer(m, method, cfl)
    T.set_velocity(fvx, fvy, cvx, cvy)
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, ny, bounds=bounds)
    ctx.transfer_configure(method, cfl)
    ctx.transfer_set_velocity(fvx, fvy, cvx, cvy)
    # a disk of material (sharp and smoothed edges)
    r = np.sqrt((cl["cx"] - 0.35) ** 2 + (cl["cy"] - 0.5) ** 2)
    u0 = np.clip((0.18 - r) * ny + 0.5, 0.0, 1.0)
    T.set_state(u0)
    ctx.transfer_set_state(u0)
    return m, T, ctx


def same(ctx, T, what, slopes=False):
    for a, b, name in zip(ctx.transfer_state(), T.state(), ("u", "nx", "ny")):
        assert np.array_equal(a, b), f"{what}: {name} differs at {np.flatnonzero(a != b)[:5]}"
    if slopes:
        for a, b in zip(ctx.transfer_slopes(), T.slopes()):
            assert np.array_equal(a, b), f"{what}: slopes"


@pytest.mark.parametrize("method", BUILT)
@pytest.mark.parametrize("split", [False, True])
def test_updates_match_oracle(oracle, zfv, method, split):
    m, T, ctx = pair(oracle, zfv, method)
    same(ctx, T, "initial normals")
    dt = T.compute_dt()
    assert ctx.transfer_compute_dt() == dt
    T.set_dt(dt)
    ctx.transfer_set_dt(dt)
    for step in range(8):
        for d in (("X", "Y") if split else ("ANY",)):
            T.update(d)
            ctx.transfer_update(d)
        same(ctx, T, f"step {step + 1}", slopes=method.startswith("MUSCL"))
    u = ctx.transfer_state()[0]
    assert 0.0 < np.sum(u) and np.any((u > 0.0) & (u < 1.0))


@pytest.mark.parametrize("method", ["CRP_N1", "MUSCL_MC_CRP", "WENO", "WENO_CRP"])
def test_open_box_and_several_updates_per_call(oracle, zfv, method):
    m, T, ctx = pair(oracle, zfv, method, nx=40, ny=32, bounds=(ZOE,) * 4,
                     field=lambda x, y: (np.full_like(x, 0.7), np.full_like(y, -0.3)))
    dt = 0.8 * T.compute_dt()
    T.set_dt(dt)
    ctx.transfer_set_dt(dt)
    for _ in range(6):
        T.update("ANY")
    ctx.transfer_update("ANY", 6)
    same(ctx, T, "after 6")
    T.set_flags()
    ctx.transfer_set_flags()
    assert np.array_equal(ctx.flags(), m.cells()["flag"])


def test_refusals(zfv):
    ctx = zfv.Context()
    ctx.mesh_rectangle(8, 8)
    for meth in ("CRP_N2", "VOF", "VOF_CRP", "MUSCLn", "MUSCLn_CRP"):
        with pytest.raises(zfv.ZfvError, match="error 4"):
            ctx.transfer_configure(meth)
    narrow = zfv.Context()   # WENO's five-cell stencils need 3 x 3 cells
    narrow.mesh_rectangle(2, 8)
    narrow.transfer_configure("WENO")
    narrow.transfer_set_velocity(np.ones(128), np.zeros(128), np.ones(16), np.zeros(16))
    narrow.transfer_set_state(np.full(16, 0.5))
    with pytest.raises(zfv.ZfvError, match="error 4"):
        narrow.transfer_update("ANY")
    ctx.transfer_configure("CRP_V3")
    with pytest.raises(zfv.ZfvError):
        ctx.transfer_update("ANY")   # no state
