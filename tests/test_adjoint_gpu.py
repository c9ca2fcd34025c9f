"""Discrete adjoint of the MUSCL + HLLC step against central finite
differences of the same device forward map.

J = integral over a window of p^2 at the final time; the initial state is a
smooth seeded density field with u = v = 0 and p = 1.  Directions are smooth
(seeded Gaussian random fields), eps = 1e-6, central differences; the bar is
a relative agreement of 1e-5, far above the FD truncation (eps^2) and
round-off (1e-16 J / eps) and far below any error of a wrong derivative term.
Kinks (limiter regions, wave-speed selections) are crossed with low
probability by such small perturbations of smooth data.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4
WINDOW = (0.3, 0.7, 0.3, 0.7)


def setup(zfv, nx, ny, limiter, bounds, seed):
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(nx, ny, bounds=bounds)
    ctx.configure(acc=2, flux="hllc", limiter=limiter, cfl=0.5, gamma=GAMMA)
    rho = 1.0 + 0.1 * problems.gaussian_random_field(nx, ny, seed, corr_len=0.1)
    return ctx, rho


def initial(rho, p=None):
    from zephyr_amd import problems
    p = np.ones_like(rho) if p is None else p
    z = np.zeros_like(rho)
    return [rho, z, z.copy(), p, problems.energy_rP(rho, p, GAMMA)]


def J_of(ctx, state, n_steps, every):
    ctx.set_state(*state)
    return ctx.adjoint_forward(n_steps, every, WINDOW)


@pytest.mark.parametrize("limiter,bounds", [("MC", (2, 2, 2, 2)), ("minmod", (1, 2, 1, 3))])
def test_adjoint_matches_central_differences(zfv, limiter, bounds):
    import torch
    from zephyr_amd.autograd import FinalPressure
    nx, ny, steps, every = 48, 40, 30, 7
    ctx, rho = setup(zfv, nx, ny, limiter, bounds, seed=3)
    r = torch.tensor(rho, requires_grad=True)
    J = FinalPressure.apply(r, ctx, steps, every, WINDOW, 1.0, GAMMA)
    J.backward()
    g = r.grad.numpy()
    assert np.isfinite(g).all() and np.abs(g).max() > 0.0
    from zephyr_amd import problems
    eps = 1e-6
    for k in range(4):
        d = problems.gaussian_random_field(nx, ny, 100 + k, corr_len=0.08)
        Jp = J_of(ctx, initial(rho + eps * d), steps, every)
        Jm = J_of(ctx, initial(rho - eps * d), steps, every)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.dot(g, d))
        assert abs(fd - ad) <= 1e-5 * max(abs(fd), abs(ad)), f"direction {k}: FD {fd} adjoint {ad}"


def test_adjoint_all_components(zfv):
    """dJ/dp and dJ/de of the initial state (e taken as an independent input)."""
    nx, ny, steps, every = 32, 32, 20, 6
    ctx, rho = setup(zfv, nx, ny, "MC", (2, 2, 2, 2), seed=9)
    st = initial(rho, 1.0 + 0.05 * np.cos(np.arange(nx * ny) * 0.01))
    st[1] = 0.05 * np.sin(np.arange(nx * ny) * 0.013)
    J0 = J_of(ctx, st, steps, every)
    lam = ctx.adjoint_backward()
    from zephyr_amd import problems
    eps = 1e-6
    for comp in (1, 2, 3, 4):   # u, v, p, e
        d = problems.gaussian_random_field(nx, ny, 40 + comp, corr_len=0.1)
        sp = [a.copy() for a in st]
        sm = [a.copy() for a in st]
        sp[comp] += eps * d
        sm[comp] -= eps * d
        fd = (J_of(ctx, sp, steps, every) - J_of(ctx, sm, steps, every)) / (2 * eps)
        ad = float(np.dot(lam[comp], d))
        assert abs(fd - ad) <= 1e-5 * max(abs(fd), abs(ad), 1e-12), f"component {comp}: FD {fd} adjoint {ad}"
    assert np.isfinite(J0)


def test_adjoint_forward_equals_plain_steps(zfv):
    nx, ny, steps = 24, 20, 9
    ctx, rho = setup(zfv, nx, ny, "MC", (2, 2, 2, 2), seed=1)
    J = J_of(ctx, initial(rho), steps, 4)
    after = ctx.state()
    ref, _ = setup(zfv, nx, ny, "MC", (2, 2, 2, 2), seed=1)
    ref.set_state(*initial(rho))
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
