"""The stream modes of the multi-material step (zfv_mm_set_overlap): 1
runs the interface reconstruction on a second stream beside the gradient,
predictor and marking; 2 also runs the composite fluxes there while the
cells without a composite face are updated, the others after them.  Every
cell goes through the same code in every mode, so each must equal the
one-stream step bit for bit -- state, dt and interface -- at both orders,
with and without direction splitting, on a uniform and an adapted mesh,
and whatever the composite-flux kernel's grid (64 workgroups forces up to
64 lanes per wavefront onto the list).
The one-stream step itself is checked against the oracle in
test_mm_gpu.py, test_mm_acc2_gpu.py and test_mm_options_gpu.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "P", "e", "T", "beta", "rhos")


def bubble(zfv, acc, split, n=96, amr=False):
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(n, n)
    ctx.mm_configure([problems.WATER, problems.AIR], 0.5)
    ctx.mm_set_accuracy(acc)
    if split:
        ctx.mm_set_splitting(split)
    i, j = np.divmod(np.arange(n * n), n)
    cx, cy = (i + 0.5) / n, (j + 0.5) / n
    ctx.mm_set_state(*problems.mm_bubble_shock(cx, cy, 1.0 / n, 7))
    if amr:
        for _ in range(2):
            ctx.mm_set_flags()
            ctx.mm_refine(2)
    return ctx


def same(a, b, what):
    za, zb = a.mm_state(), b.mm_state()
    for k in FIELDS:
        assert np.array_equal(za[k], zb[k], equal_nan=True), f"{what}: {k}"
    assert a.mm_dt() == b.mm_dt(), f"{what}: dt"


@pytest.mark.parametrize("acc,split,amr", [(1, None, False), (2, None, False), (1, "STRANG", False),
                                           (2, "SIMPLE", False), (1, None, True), (2, None, True)])
def test_overlapped_step_equals_one_stream(zfv, acc, split, amr):
    ref = bubble(zfv, acc, split, amr=amr)
    ref.mm_set_overlap(0)
    forms = [(1, 0), (2, 0), (2, 64), (1, 512)]
    runs = []
    for mode, groups in forms:
        c = bubble(zfv, acc, split, amr=amr)
        c.mm_set_overlap(mode, groups)
        runs.append(c)
    for step in range(10):
        for c in [ref] + runs:
            c.mm_step(1)
        for f, c in zip(forms, runs):
            same(c, ref, f"step {step + 1}, mode / groups {f}")
    assert ref.mm_crp_faces() > 0, "the shock-bubble has composite faces"
    ia, ib = ref.mm_interface(), runs[0].mm_interface()
    for a, b in zip(ia, ib):
        assert np.array_equal(a, b)


def test_overlap_switches_between_steps(zfv):
    """Switching the form between steps (and several steps per call) keeps
    the results; the AMR cycle (flags, refine, steps) runs overlapped."""
    a, b = bubble(zfv, 1, None), bubble(zfv, 1, None)
    a.mm_set_overlap(0)
    for k, mode in enumerate((2, 0, 1, 2)):
        b.mm_set_overlap(mode)
        a.mm_step(3)
        b.mm_step(3)
        same(a, b, f"call {k}")
        for c in (a, b):
            c.mm_set_flags()
            c.mm_refine(2)
        assert a.n == b.n
    for bad in ((1, -1), (3, 0)):
        with pytest.raises(zfv.ZfvError):
            b.mm_set_overlap(*bad)
