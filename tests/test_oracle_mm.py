"""Known answers for the oracle's multi-material restatement (MmFluid, first
order, no composite Riemann problems; MixturePT closure; PLIC formulas).

The reference's own tests pin none of this path, so each part is pinned by
what the mathematics fixes:
* PLIC: a line section and its volume fraction are inverse maps; the face
  fraction of two cells cut by one straight line is the line's height on the
  shared face (brute-force geometry); the average flux at CFL 1 is the cell
  fraction.
* Closure: pure materials in closed form; a mixture's (P, T) satisfies
  1/rho = sum beta_i v_i(P, T), e = sum beta_i e_i(P, T); (rho, P) -> e ->
  (rho, e) round trips; the mixture sound speed equals dP/drho along the
  isentrope de = P / rho^2 drho (central differences); a mixture of two
  identical materials behaves as one.
* Flux: HLLC consistency F(z, z) = Flux(z); for one ideal gas the mmf HLLC is
  the single-material HLLC bit for bit.
* Step: one material reproduces the single-material first-order step bit for
  bit (same dt through max_dt); periodic boxes conserve mass, partial
  masses, momentum and energy; a uniform mixture stays uniform.
"""
import numpy as np
import pytest

WALL, ZOE, PERIODIC = 1, 2, 3   # Boundary ids (geom/boundary.h:9-16)
WATER = (4.4, 6.0e8, 0.0, 4160.0, 230.7, 1000.0)   # StiffenedGas("Water") (stiffened_gas.cpp:33-40)
AIR = (1.4, 0.0, 0.0, 718.0, 0.0, 1.25)            # StiffenedGas("Air") (stiffened_gas.cpp:22-25)
GAS = (1.4, 0.0, 0.0, 2.5, 0.0, 1.0)               # StiffenedGas("Gas")


def test_section_and_volume_fraction_are_inverse(oracle):
    rng = np.random.default_rng(0)
    for _ in range(2000):
        a = rng.uniform(0.0, 1.0)
        th = rng.uniform(0.0, 2.0 * np.pi)
        A, B = rng.uniform(0.3, 3.0, 2)
        p = oracle.quad_find_section(a, np.cos(th), np.sin(th), A, B)
        assert abs(oracle.quad_volume_fraction(p, np.cos(th), np.sin(th), A, B) - a) < 1e-14
        # complement: the section of 1 - alpha is the mirrored line
        q = oracle.quad_find_section(1.0 - a, np.cos(th), np.sin(th), A, B)
        assert abs(p + q) < 1e-12 * (A + B)


def _area_below(k, m, x0, x1, n=200001):
    xs = np.linspace(x0, x1, n)
    return np.trapezoid(np.clip(k * xs + m, 0.0, 1.0), xs)


def test_face_fraction_of_a_straight_interface(oracle):
    """Two unit cells [-1,0]x[0,1] and [0,1]x[0,1] cut by y = k x + m: the
    PLIC face fraction of their volume fractions is the line's height m on
    the shared face x = 0."""
    rng = np.random.default_rng(1)
    n = 0
    for _ in range(600):
        m = rng.uniform(0.02, 0.98)
        k = np.tan(rng.uniform(-1.4, 1.4))
        a1, a2 = _area_below(k, m, -1.0, 0.0), _area_below(k, m, 0.0, 1.0)
        if min(a1, a2) < 1e-6 or max(a1, a2) > 1.0 - 1e-6:
            continue
        assert abs(oracle.face_fraction(a1, a2) - m) < 1e-7
        n += 1
    assert n > 400


def test_face_fraction_variants(oracle):
    rng = np.random.default_rng(2)
    for _ in range(500):
        a1, a2 = rng.uniform(0, 1, 2)
        for kind in ("plic", "s", "v3", "v5"):
            f = oracle.face_fraction(a1, a2, kind)
            assert f == oracle.face_fraction(a2, a1, kind)          # symmetric
            assert min(a1, a2) - 1e-15 <= f <= max(a1, a2) + 1e-15 or kind in ("v3", "v5")
        assert oracle.face_fraction(a1, a1) == a1
    assert oracle.face_fraction(0.0, 1.0) == 0.5
    assert oracle.face_fraction(0.0, 0.4, "v5") == 0.0 and oracle.face_fraction(0.3, 1.0, "v3") == 1.0


def test_average_flux_at_cfl_one_is_the_cell_fraction(oracle):
    rng = np.random.default_rng(3)
    for _ in range(1000):
        a, c = rng.uniform(0.01, 0.99), rng.uniform(-0.99, 0.99)
        assert abs(oracle.average_flux(a, c, 1.0) - a) < 1e-14


def test_pure_closure_closed_forms(oracle):
    g, P0, e0, Cv, T0, _ = WATER
    rho, P = 1000.0, 1.0e5
    e, T, c, rh = oracle.mm_closure_rP([WATER, AIR], rho, P, [1.0, 0.0])
    assert e == e0 + (P + g * P0) / ((g - 1.0) * rho)
    assert T == T0 + (P + P0) / ((g - 1.0) * Cv * rho)
    assert c == np.sqrt(g * (P + P0) / rho)
    assert rh[0] == rho and np.isnan(rh[1])
    P2, T2, _ = oracle.mm_closure([WATER, AIR], rho, e, [1.0, 0.0])
    assert P2 == (g - 1.0) * rho * (e - e0) - g * P0 and T2 == T0 + (P2 + P0) / ((g - 1.0) * Cv * rho)


def _v_e(mat, P, T):
    g, P0, e0, Cv, T0, _ = mat
    return (g - 1.0) * Cv * (T - T0) / (P + P0), e0 + (P + g * P0) / (P + P0) * Cv * (T - T0)


@pytest.mark.parametrize("beta", [[0.5, 0.5], [0.999, 0.001], [0.01, 0.99], [0.3, 0.7]])
def test_mixture_closure_equilibrium(oracle, beta):
    mats = [WATER, AIR]
    rho = 1.0 / (beta[0] / 1000.0 + beta[1] / 1.2)
    P = 2.0e5
    e, T, c, rh = oracle.mm_closure_rP(mats, rho, P, beta)
    v = [_v_e(m, P, T)[0] for m in mats]
    em = [_v_e(m, P, T)[1] for m in mats]
    assert abs(beta[0] * v[0] + beta[1] * v[1] - 1.0 / rho) < 1e-11 / rho
    assert abs(beta[0] * em[0] + beta[1] * em[1] - e) < 1e-12 * abs(e)
    assert np.allclose(rh[:2], [1.0 / v[0], 1.0 / v[1]], rtol=1e-11)
    # (rho, e) -> (P, T) returns the same equilibrium
    P2, T2, rh2 = oracle.mm_closure(mats, rho, e, beta, 1.0e5, 300.0)
    assert abs(P2 - P) < 1e-9 * P and abs(T2 - T) < 1e-9 * T
    # sound speed: dP/drho along de = P / rho^2 drho
    h = 1e-6 * rho
    Pp, _, _ = oracle.mm_closure(mats, rho + h, e + P / rho ** 2 * h, beta, P, T)
    Pm, _, _ = oracle.mm_closure(mats, rho - h, e - P / rho ** 2 * h, beta, P, T)
    c2 = (Pp - Pm) / (2.0 * h)
    assert abs(c2 - c * c) < 1e-5 * c * c


def test_identical_materials_mix_as_one(oracle):
    rho, P = 1.3, 2.5
    e1, T1, c1, _ = oracle.mm_closure_rP([GAS, GAS], rho, P, [1.0, 0.0])
    e2, T2, c2, _ = oracle.mm_closure_rP([GAS, GAS], rho, P, [0.4, 0.6])
    assert abs(e2 - e1) < 1e-12 * e1 and abs(T2 - T1) < 1e-12 * T1 and abs(c2 - c1) < 1e-10 * c1


def _mesh_state(oracle, nx, ny, bounds, mats, seed, mixed=True):
    m = oracle.Mesh.rectangle(nx, ny, bounds=bounds)
    cl = m.cells()
    cx, cy = cl["cx"], cl["cy"]
    rng = np.random.default_rng(seed)
    n = m.n
    r = 1.0 + 0.2 * np.sin(3 * cx) * np.cos(2 * cy) + 0.05 * rng.random(n)
    u = 0.3 * np.cos(2 * cy) + 0.02 * rng.random(n)
    v = -0.2 * np.sin(3 * cx) + 0.02 * rng.random(n)
    p = 1.0 + 0.3 * np.cos(4 * cx + cy)
    b0 = np.clip(0.5 + 0.6 * np.sin(5 * cx + 2 * cy), 0.0, 1.0) if mixed else np.ones(n)
    beta = np.stack([b0, 1.0 - b0]) if len(mats) == 2 else b0[None, :]
    return m, (r, u, v, p, beta)


def test_hllc_consistency(oracle):
    """F(z, z) = Flux(z) for pure and mixed states (hllc.cpp:339-421)."""
    m, (r, u, v, p, beta) = _mesh_state(oracle, 8, 6, (ZOE,) * 4, [GAS, GAS], 11)
    mats = [GAS, (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)]
    mm = oracle.MmFluid(m, mats)
    mm.set_state(r, u, v, p, beta)
    z = mm.state()
    for c in range(m.n):
        F = mm.face_flux(c, c, 1.0, 0.0)
        ru = z["rho"][c] * z["u"][c]
        E = z["e"][c] + 0.5 * (z["u"][c] ** 2 + z["v"][c] ** 2)
        want = [ru, ru * z["u"][c] + z["P"][c], ru * z["v"][c], 0.0, (z["rho"][c] * E + z["P"][c]) * z["u"][c],
                ru * z["beta"][0][c], ru * z["beta"][1][c], 0.0, 0.0]
        assert np.allclose(F, want, rtol=1e-13, atol=1e-15), c


def test_one_material_equals_single_fluid_step(oracle):
    """MmFluid with one ideal gas takes the SmFluid first-order HLLC step bit
    for bit when both use the same dt (max_dt below both CFL limits)."""
    m, (r, u, v, p, _) = _mesh_state(oracle, 24, 18, (WALL, ZOE, PERIODIC, PERIODIC),
                                     [GAS], 5, mixed=False)
    mm = oracle.MmFluid(m, [GAS], cfl=0.5, crp=True)
    mm.set_state(r, u, v, p, np.ones((1, m.n)))
    sm = oracle.SmFluid(m, gamma=1.4, flux=oracle.HLLC, acc=1, cfl=0.5)
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


@pytest.mark.parametrize("crp", [False, True])
def test_periodic_box_conserves_mass_momentum_energy(oracle, crp):
    """Without composite Riemann problems also every partial mass; with them
    the closure's cutoff of mass fractions below 1e-12 (including negative
    ones a composite flux can leave, models.cpp:195-196) moves partial mass
    between materials, while the totals stay conserved."""
    m, st = _mesh_state(oracle, 20, 16, (PERIODIC,) * 4, [WATER, AIR], 7)
    r, u, v, p, beta = st
    mm = oracle.MmFluid(m, [GAS, (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)], cfl=0.5, crp=crp)   # Gas / HeavyGas
    mm.set_state(r, u, v, p, beta)
    vol = m.cells()["volume"]

    def totals():
        z = mm.state()
        E = z["rho"] * (z["e"] + 0.5 * (z["u"] ** 2 + z["v"] ** 2))
        return np.array([np.sum(vol * z["rho"]), np.sum(vol * z["rho"] * z["u"]), np.sum(vol * z["rho"] * z["v"]),
                         np.sum(vol * E), *(np.sum(vol * z["rho"] * z["beta"][i]) for i in range(2))])

    t0 = totals()
    for _ in range(10):
        assert mm.update() == 0
    t1 = totals()
    k = 6 if not crp else 4
    assert np.all(np.abs(t1 - t0)[:k] <= 1e-12 * np.maximum(np.abs(t0), 1.0)[:k])


def test_plic_composite_fluxes_keep_an_advected_interface_sharp(oracle):
    """A heavy-gas disc advected diagonally through a periodic box: with the
    PLIC composite Riemann fluxes (the MmFluid default) the interface stays
    a few cells wide; plain HLLC smears it over many cells."""
    heavy = (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)
    counts = {}
    for crp in (False, True):
        m = oracle.Mesh.rectangle(40, 40, bounds=(PERIODIC,) * 4)
        cl = m.cells()
        inside = ((cl["cx"] - 0.5) ** 2 + (cl["cy"] - 0.5) ** 2 < 0.2 ** 2).astype(float)
        mm = oracle.MmFluid(m, [GAS, heavy], cfl=0.5, crp=crp)
        mm.set_state(np.where(inside > 0, 4.0, 1.0), np.full(m.n, 1.0), np.full(m.n, 0.5), np.ones(m.n),
                     np.stack([1.0 - inside, inside]))
        for _ in range(60):
            assert mm.update() == 0
        a = mm.state()["beta"][1]
        counts[crp] = int(np.sum((a > 1e-6) & (a < 1.0 - 1e-6)))
    assert counts[True] < 0.2 * counts[False], counts


def test_uniform_mixture_stays_uniform(oracle):
    m = oracle.Mesh.rectangle(12, 10, bounds=(ZOE,) * 4)
    n = m.n
    mats = [GAS, (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)]
    mm = oracle.MmFluid(m, mats, cfl=0.5)
    mm.set_state(np.full(n, 2.0), np.full(n, 0.3), np.full(n, -0.1), np.full(n, 1.5),
                 np.stack([np.full(n, 0.3), np.full(n, 0.7)]))
    z0 = mm.state()
    for _ in range(5):
        assert mm.update() == 0
    z1 = mm.state()
    for k in ("rho", "u", "v", "P"):
        assert np.allclose(z1[k], z0[k], rtol=1e-13, atol=1e-14), k


def test_interface_normal_of_a_vertical_and_a_horizontal_front(oracle):
    """A column (row) of half-filled cells between pure cells: the PLIC
    normal of the first material points across the front and the section
    passes through the cell centre."""
    nx = ny = 8
    m = oracle.Mesh.rectangle(nx, ny, bounds=(ZOE,) * 4)
    cl = m.cells()
    cx, cy = cl["cx"], cl["cy"]
    mats = [GAS, GAS]
    for axis in (0, 1):
        coord = cx if axis == 0 else cy
        b0 = np.where(coord < 0.375, 1.0, np.where(coord < 0.5, 0.5, 0.0))
        mm = oracle.MmFluid(m, mats)
        mm.set_state(np.ones(m.n), np.zeros(m.n), np.zeros(m.n), np.ones(m.n), np.stack([b0, 1.0 - b0]))
        nxs, nys, ps = mm.interface()
        mixed = (b0 > 0.0) & (b0 < 1.0)
        want = (1.0, 0.0) if axis == 0 else (0.0, 1.0)
        assert np.allclose(nxs[0][mixed], want[0]) and np.allclose(nys[0][mixed], want[1])
        assert np.allclose(ps[0][mixed], 0.0, atol=1e-15)
        assert np.all(nxs[0][~mixed] == 0.0) and np.all(ps[0][~mixed] == 0.0)
