"""Discrete adjoint of the MUSCL + HLLC step against central finite
differences of the same device forward map.

J = integral over a window of p^2 at the final time.  The initial state is a
seeded smooth flow in which every component varies (rho, p = 1 + 0.1 g,
u, v = 0.1 g): a component that is identically zero puts every limiter
product psi(a, b) = phi(a / b) b at its corner (0, 0), where psi is not
differentiable and a central quotient measures psi(d) instead of a linear
function of d.

The step is piecewise smooth: the limiters and the HLLC wave-speed selection
switch branches.  With eps = 1e-7 a perturbation crosses a switch with low
probability; rounding contributes about 1e-16 J / eps ~ 1e-10.  Bars: 1e-6
relative for the smooth slopes (limiter none / unit), 1e-5 for minmod and
MC, both far below what a dropped or wrong derivative term produces
(1e-3 relative and more).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4
WINDOW = (0.3, 0.7, 0.3, 0.7)
EPS = 1e-7


def setup(zfv, nx, ny, limiter, bounds, seed):
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, ny, bounds=bounds)
    ctx.configure(acc=2, flux="hllc", limiter=limiter, cfl=0.5, gamma=GAMMA)
    return ctx, list(problems.smooth_flow_field(nx, ny, seed, GAMMA, corr_len=0.1))


def J_of(ctx, state, n_steps, every):
    ctx.set_state(*state)
    return ctx.adjoint_forward(n_steps, every, WINDOW)


def with_rho(state, rho):
    from zephyr_amd import problems
    return [rho, state[1], state[2], state[3], problems.energy_rP(rho, state[3], GAMMA)]


@pytest.mark.parametrize("limiter,bounds,tol", [
    ("none", (2, 2, 2, 2), 1e-6), ("unit", (1, 2, 3, 3), 1e-6),
    ("MC", (2, 2, 2, 2), 1e-5), ("minmod", (1, 2, 3, 3), 1e-5), ("superbee", (3, 3, 1, 1), 1e-5),
])
def test_density_gradient_matches_central_differences(zfv, limiter, bounds, tol):
    """dJ/drho0 with p, u, v held fixed, through the torch.autograd binding
    (chain rule through e = p / ((gamma - 1) rho))."""
    import torch
    from zephyr_amd import problems
    from zephyr_amd.autograd import FinalPressure
    nx, ny, steps, every = 48, 40, 30, 7
    ctx, st = setup(zfv, nx, ny, limiter, bounds, seed=3)
    r = torch.tensor(st[0], requires_grad=True)
    J = FinalPressure.apply(r, ctx, steps, every, WINDOW, st[3], GAMMA, st[1], st[2])
    J.backward()
    g = r.grad.numpy()
    assert np.isfinite(g).all() and np.abs(g).max() > 0.0
    for k in range(4):
        d = problems.gaussian_random_field(nx, ny, 100 + k, corr_len=0.08)
        Jp = J_of(ctx, with_rho(st, st[0] + EPS * d), steps, every)
        Jm = J_of(ctx, with_rho(st, st[0] - EPS * d), steps, every)
        fd = (Jp - Jm) / (2 * EPS)
        ad = float(np.dot(g, d))
        assert abs(fd) > 1e-4, f"direction {k}: J barely depends on rho0 ({fd})"
        assert abs(fd - ad) <= tol * max(abs(fd), abs(ad)), f"direction {k}: FD {fd} adjoint {ad}"


@pytest.mark.parametrize("limiter", ["MC", "unit"])
def test_adjoint_all_components(zfv, limiter):
    """dJ/d(rho, u, v, p, e) of the initial state, e an independent input."""
    from zephyr_amd import problems
    nx, ny, steps, every = 32, 32, 20, 6
    ctx, st = setup(zfv, nx, ny, limiter, (2, 2, 2, 2), seed=9)
    J0 = J_of(ctx, st, steps, every)
    lam = ctx.adjoint_backward()
    for comp in range(5):
        d = problems.gaussian_random_field(nx, ny, 40 + comp, corr_len=0.1)
        sp = [a.copy() for a in st]
        sm = [a.copy() for a in st]
        sp[comp] += EPS * d
        sm[comp] -= EPS * d
        fd = (J_of(ctx, sp, steps, every) - J_of(ctx, sm, steps, every)) / (2 * EPS)
        ad = float(np.dot(lam[comp], d))
        assert abs(fd - ad) <= 1e-5 * max(abs(fd), abs(ad), 1e-3), f"component {comp}: FD {fd} adjoint {ad}"
    assert np.isfinite(J0)


def test_adjoint_forward_equals_plain_steps(zfv):
    nx, ny, steps = 24, 20, 9
    ctx, st = setup(zfv, nx, ny, "MC", (2, 2, 2, 2), seed=1)
    J = J_of(ctx, st, steps, 4)
    after = ctx.state()
    ref, _ = setup(zfv, nx, ny, "MC", (2, 2, 2, 2), seed=1)
    ref.set_state(*st)
    ref.step(steps)
    for a, b in zip(after, ref.state()):
        assert np.array_equal(a, b)
    cells, _, _ = ref.mesh_download()
    x, y, V = cells["cx"], cells["cy"], cells["volume"]
    p = ref.state()[3]
    inside = (x >= WINDOW[0]) & (x <= WINDOW[1]) & (y >= WINDOW[2]) & (y <= WINDOW[3])
    Jr = 0.0
    for c in np.flatnonzero(inside):
        Jr = Jr + p[c] * p[c] * V[c]
    assert J == Jr
    # backward leaves the final state in place
    ctx.adjoint_backward()
    for a, b in zip(ctx.state(), ref.state()):
        assert np.array_equal(a, b)


def test_unpaired_periodic_side_is_refused(zfv):
    """A periodic side whose opposite side is not periodic links faces that
    are seen from one side only; the reverse gather cannot serve it."""
    ctx, st = setup(zfv, 16, 12, "MC", (1, 2, 1, 3), seed=2)
    ctx.set_state(*st)
    with pytest.raises(zfv.ZfvError, match="unpaired periodic"):
        ctx.adjoint_forward(3, 1, WINDOW)


def test_backward_is_repeatable_and_checkpoint_independent(zfv):
    """The reverse sweep recomputes segments from checkpoints: the gradient
    must not depend on the checkpoint interval beyond the reordering of the
    dt-adjoint sums (atomic adds)."""
    nx, ny, steps = 32, 24, 12
    ctx, st = setup(zfv, nx, ny, "MC", (2, 2, 2, 2), seed=5)
    grads = []
    for every in (1, 5, 12):
        J_of(ctx, st, steps, every)
        grads.append(ctx.adjoint_backward())
    for g in grads[1:]:
        for a, b in zip(grads[0], g):
            assert np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(grads[0][0]).max())


def test_checkpoint_list_growth_keeps_results(zfv):
    """A context reused with more checkpoints than before (the checkpoint
    list reallocates) gives the gradient of a fresh context, and a later run
    with fewer checkpoints still works."""
    nx, ny = 24, 20
    ctx, st = setup(zfv, nx, ny, "MC", (2, 2, 2, 2), seed=4)
    for steps, every in ((4, 2), (13, 1), (13, 13)):
        Ja = J_of(ctx, st, steps, every)
        ga = ctx.adjoint_backward()
        fresh, _ = setup(zfv, nx, ny, "MC", (2, 2, 2, 2), seed=4)
        Jb = J_of(fresh, st, steps, every)
        gb = fresh.adjoint_backward()
        assert Ja == Jb
        for a, b in zip(ga, gb):
            assert np.allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(gb[0]).max())
