"""Device MmFluid options against the oracle on the same inputs: the plain
mmf fluxes of set_method (HLL, Rusanov), the heuristic CrpModes V3 / V5 /
VS and direction splitting (SIMPLE, STRANG), first and second order.

Bar: bit-exact state (NaN where a material is absent) and dt after every
step, on power-of-two bases (see test_mm_amr_gpu).
"""
import numpy as np
import pytest

from test_mm_amr_gpu import bubble, compare_mesh
from test_mm_gpu import GAS, HEAVY, WATER, AIR, assert_same

pytestmark = pytest.mark.gpu

THIRD = (1.67, 0.0, 0.0, 1.5, 0.0, 0.5)


def configured_pair(oracle, zfv, mats, acc, method=None, crp_mode=None, split=None, bounds=(2, 2, 2, 2), nx=32):
    m = oracle.Mesh.rectangle(nx, nx, bounds=bounds)
    o = oracle.MmFluid(m, mats, 0.5)
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, nx, bounds=bounds)
    ctx.mm_configure(mats, 0.5)
    for side in (o, ctx):
        acc_set = side.set_accuracy if side is o else side.mm_set_accuracy
        acc_set(acc)
    if method:
        o.set_method(method)
        ctx.mm_set_method(method)
    if crp_mode:
        o.set_crp_mode(crp_mode)
        ctx.mm_set_crp_mode(crp_mode)
    if split:
        o.set_splitting(split)
        ctx.mm_set_splitting(split)
    cl = m.cells()
    st = list(bubble(cl["cx"], cl["cy"], mats[:2], 5))
    if len(mats) == 3:
        b = st[4]
        band = np.clip(1.0 - np.abs(cl["cy"] - 0.2) / 0.08, 0.0, 1.0)
        st[4] = np.stack([b[0] * (1.0 - band), b[1] * (1.0 - band), band])
    o.set_state(*st)
    ctx.mm_set_state(*st)
    return m, o, ctx


def run_and_compare(o, ctx, steps, what):
    for step in range(steps):
        assert o.update() == 0, f"{what} step {step}: the oracle rejected a state"
        ctx.mm_step(1)
        assert ctx.mm_dt() == o.dt(), f"{what} step {step}: dt"
        assert_same(ctx.mm_state(), o.state(), f"{what} step {step}")


@pytest.mark.parametrize("acc", [1, 2])
@pytest.mark.parametrize("method", ["HLL", "Rusanov"])
@pytest.mark.parametrize("mats", [[GAS, HEAVY], [GAS, HEAVY, THIRD]])
def test_plain_mmf_fluxes_match_oracle(oracle, zfv, method, acc, mats):
    # no periodic sides: at second order the reference extrapolates a periodic
    # face's neighbour state from that neighbour's own centre, across the
    # domain, which these plain fluxes do not survive
    _, o, ctx = configured_pair(oracle, zfv, mats, acc, method=method, bounds=(1, 2, 2, 1))
    run_and_compare(o, ctx, 6, method)
    assert ctx.mm_crp_faces() == 0


@pytest.mark.parametrize("acc", [1, 2])
@pytest.mark.parametrize("mode", ["V3", "V5", "VS"])
@pytest.mark.parametrize("mats", [[GAS, HEAVY], [WATER, AIR]])
def test_heuristic_crp_modes_match_oracle(oracle, zfv, mode, acc, mats):
    _, o, ctx = configured_pair(oracle, zfv, mats, acc, crp_mode=mode, bounds=(2, 1, 2, 2))
    run_and_compare(o, ctx, 6, mode)
    assert ctx.mm_crp_faces() > 0


@pytest.mark.parametrize("acc", [1, 2])
@pytest.mark.parametrize("split", ["SIMPLE", "STRANG"])
@pytest.mark.parametrize("method,mode", [(None, None), ("HLL", None), (None, "V5")])
def test_direction_splitting_matches_oracle(oracle, zfv, split, acc, method, mode):
    _, o, ctx = configured_pair(oracle, zfv, [GAS, HEAVY], acc, method=method, crp_mode=mode, split=split,
                                bounds=(1, 2, 2, 1))
    run_and_compare(o, ctx, 6, split)


def test_splitting_on_an_adapted_mesh(oracle, zfv):
    m, o, ctx = configured_pair(oracle, zfv, [WATER, AIR], 2, split="STRANG")
    for cycle in range(3):
        o.set_flags()
        ctx.mm_set_flags()
        assert o.refine(2) >= 0
        ctx.mm_refine(2)
        compare_mesh(ctx, m, f"cycle {cycle}")
        run_and_compare(o, ctx, 3, f"cycle {cycle}")
    assert np.max(m.cells()["level"]) == 2


@pytest.mark.skip(reason="the device multi-material Godunov flux is not built: its second-order step "
                         "made an illegal memory access on the device whose cause was not found; the oracle's "
                         "flux stays pinned in test_oracle_mm_godunov.py")
@pytest.mark.parametrize("acc", [1, 2])
@pytest.mark.parametrize("mats", [[GAS, HEAVY], [WATER, AIR]])
def test_mmf_godunov_matches_oracle(oracle, zfv, acc, mats):
    """set_method(Godunov): the exact Riemann solver takes pow in its
    rarefaction and shock functions, whose device and glibc values may
    differ in the last bit, so the bar is relative 1e-10 per field after 6
    steps (the same as the single-material Godunov step's 1e-12 over 30
    steps, with the closure's Newton iterations in between)."""
    _, o, ctx = configured_pair(oracle, zfv, mats, acc, method="Godunov", bounds=(1, 2, 2, 1))
    for step in range(6):
        assert o.update() == 0, f"step {step}: the oracle rejected a state"
        ctx.mm_step(1)
        np.testing.assert_allclose(ctx.mm_dt(), o.dt(), rtol=1e-10)
    zg, zo = ctx.mm_state(), o.state()
    for k in ("rho", "u", "v", "P", "e", "T", "beta"):
        a, b = np.nan_to_num(zg[k]), np.nan_to_num(zo[k])
        np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-10 * np.abs(b).max(), err_msg=k)
    assert ctx.mm_crp_faces() == 0


def test_unsupported_options_are_refused(zfv):
    ctx = zfv.Context()
    ctx.mesh_rectangle(8, 8)
    ctx.mm_configure([GAS, HEAVY], 0.5)
    for bad in (lambda: ctx.mm_set_method("Godunov"), lambda: ctx.mm_set_method(4),
                lambda: ctx.mm_set_crp_mode("MUSCL")):
        with pytest.raises(zfv.ZfvError, match="error 4"):
            bad()
    ctx.mm_set_crp_mode("NONE")   # ignored, as MmFluid::set_crp_mode does
