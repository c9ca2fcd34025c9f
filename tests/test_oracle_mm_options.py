"""Known answers for the oracle's MmFluid options: the plain mmf fluxes of
set_method (HLL, Rusanov: hll.cpp:125-160, rusanov.cpp:46-83), the
heuristic CrpModes V3 / V5 / VS (alpha_sigma, mm_fluid.cpp:228-256) and
direction splitting (update / integrate with EuCell::faces(dir),
mm_fluid.cpp:112-170, amr_faces.cpp:186-197).

Pins:
* one ideal gas with Rusanov reproduces the single-fluid step bit for bit
  (same dt through max_dt), first and second order; with HLL, whose mmf
  wave-speed bounds differ from the single-fluid ones, the first step at a
  uniform sound speed, where the bounds agree;
* a periodic box conserves mass, momentum, energy and every partial mass
  with the plain fluxes, and mass, momentum and energy with each
  composite-flux mode; every mode keeps an advected interface sharper than
  no composite fluxes;
* for data that vary along x only, the Y sweep changes nothing, so SIMPLE
  equals the unsplit step and STRANG with dt equals two unsplit steps of
  dt / 2 (to round-off: each sweep re-closes the state from its conserved
  vector); for data that vary along y only the X sweeps change nothing, and
  STRANG equals one unsplit step.
"""
import numpy as np
import pytest

from test_oracle_mm import GAS, PERIODIC, WALL, ZOE, _mesh_state

HEAVY = (1.1, 0.0, 0.0, 2.5, 0.0, 4.0)


def _single_fluid_pair(oracle, flux, acc, r, u, v, p, m):
    mm = oracle.MmFluid(m, [GAS], cfl=0.5)
    mm.set_method(flux)
    mm.set_accuracy(acc)
    mm.set_state(r, u, v, p, np.ones((1, m.n)))
    sm = oracle.SmFluid(m, gamma=1.4, flux={"HLL": oracle.HLL, "Rusanov": oracle.RUSANOV}[flux], acc=acc, cfl=0.5)
    z = mm.state()
    sm.set_state(z["rho"], z["u"], z["v"], z["P"], z["e"])
    for st in (mm, sm):
        st.set_max_dt(1e-3)
    return mm, sm


def _assert_same_step(mm, sm, step):
    assert mm.update() == 0
    sm.update(nthreads=1)
    zm, zs = mm.state(), sm.state()
    for k, a in zip(("rho", "u", "v", "P", "e"), zs):
        assert np.array_equal(zm[k], a), f"{k} at step {step}"


@pytest.mark.parametrize("acc", [1, 2])
def test_one_material_rusanov_equals_single_fluid(oracle, acc):
    m, (r, u, v, p, _) = _mesh_state(oracle, 24, 18, (WALL, ZOE, ZOE, WALL), [GAS], 5, mixed=False)
    mm, sm = _single_fluid_pair(oracle, "Rusanov", acc, r, u, v, p, m)
    for step in range(5):
        _assert_same_step(mm, sm, step)


def test_one_material_hll_equals_single_fluid_at_uniform_sound_speed(oracle):
    """The mmf HLL bounds S_L with c_L on both sides and S_R with c_R
    (hll.cpp:131-138), the single-fluid HLL with each side's own speed and
    zero; with one sound speed everywhere and subsonic flow the two agree,
    so the first first-order step is the same bit for bit (second-order
    face states have their own sound speeds).  Blocky power-of-two
    densities with P = rho make c = sqrt(1.4 P / rho) exactly uniform."""
    m = oracle.Mesh.rectangle(24, 18, bounds=(WALL, ZOE, ZOE, WALL))
    cl = m.cells()
    r = np.where(cl["cx"] < 0.4, 1.0, 2.0) * np.where(cl["cy"] < 0.6, 1.0, 0.5)
    rng = np.random.default_rng(3)
    u, v = 0.3 * rng.standard_normal(m.n), 0.3 * rng.standard_normal(m.n)
    mm, sm = _single_fluid_pair(oracle, "HLL", 1, r, u, v, r.copy(), m)
    _assert_same_step(mm, sm, 0)


def _totals(m, mm):
    z = mm.state()
    vol = m.cells()["volume"]
    E = z["rho"] * (z["e"] + 0.5 * (z["u"] ** 2 + z["v"] ** 2))
    return np.array([np.sum(vol * z["rho"]), np.sum(vol * z["rho"] * z["u"]), np.sum(vol * z["rho"] * z["v"]),
                     np.sum(vol * E), *(np.sum(vol * z["rho"] * z["beta"][i]) for i in range(2))])


@pytest.mark.parametrize("setup", [("method", "HLL"), ("method", "Rusanov"), ("crp", "V3"), ("crp", "V5"),
                                   ("crp", "VS")])
def test_periodic_box_conserves(oracle, setup):
    m, (r, u, v, p, beta) = _mesh_state(oracle, 20, 16, (PERIODIC,) * 4, [GAS, HEAVY], 7)
    mm = oracle.MmFluid(m, [GAS, HEAVY], cfl=0.5)
    if setup[0] == "method":
        mm.set_method(setup[1])
    else:
        mm.set_crp_mode(setup[1])
    mm.set_state(r, u, v, p, beta)
    t0 = _totals(m, mm)
    for _ in range(10):
        assert mm.update() == 0
    t1 = _totals(m, mm)
    k = 6 if setup[0] == "method" else 4
    assert np.all(np.abs(t1 - t0)[:k] <= 1e-12 * np.maximum(np.abs(t0), 1.0)[:k])


def _advected_width(oracle, mode):
    m = oracle.Mesh.rectangle(40, 40, bounds=(PERIODIC,) * 4)
    cl = m.cells()
    inside = ((cl["cx"] - 0.5) ** 2 + (cl["cy"] - 0.5) ** 2 < 0.2 ** 2).astype(float)
    mm = oracle.MmFluid(m, [GAS, HEAVY], cfl=0.5, crp=mode != "NONE")
    if mode != "NONE":
        mm.set_crp_mode(mode)
    mm.set_state(np.where(inside > 0, 4.0, 1.0), np.full(m.n, 1.0), np.full(m.n, 0.5), np.ones(m.n),
                 np.stack([1.0 - inside, inside]))
    for _ in range(60):
        assert mm.update() == 0
    a = mm.state()["beta"][1]
    return int(np.sum((a > 1e-6) & (a < 1.0 - 1e-6)))


def test_heuristic_composite_modes_keep_an_interface_sharper(oracle):
    smeared = _advected_width(oracle, "NONE")
    widths = {mode: _advected_width(oracle, mode) for mode in ("V3", "V5", "VS", "PLIC")}
    assert all(w < 0.5 * smeared for w in widths.values()), (smeared, widths)
    assert len(set(widths.values())) > 1, widths   # the modes differ


def _one_d(oracle, along_x, split, acc, crp, dt, steps):
    m = oracle.Mesh.rectangle(32, 32, bounds=(ZOE, ZOE, WALL, WALL) if along_x else (WALL, WALL, ZOE, ZOE))
    cl = m.cells()
    s = cl["cx"] if along_x else cl["cy"]
    a = np.clip(1.0 - np.abs(s - 0.45) / 0.12, 0.0, 1.0)
    r = 1.0 + 3.0 * a
    vel = 0.3 + 0.1 * np.sin(5 * s)
    z0 = np.zeros(m.n)
    mm = oracle.MmFluid(m, [GAS, HEAVY], cfl=0.5, crp=crp)
    mm.set_accuracy(acc)
    mm.set_splitting(split)
    mm.set_max_dt(dt)
    mm.set_state(r, vel if along_x else z0, z0 if along_x else vel, 1.0 + 0.2 * np.cos(3 * s),
                 np.stack([1.0 - a * 4.0 / r, a * 4.0 / r]))
    for _ in range(steps):
        assert mm.update() == 0
    return mm.state()


@pytest.mark.parametrize("along_x", [True, False])
@pytest.mark.parametrize("acc,crp", [(1, True), (1, False), (2, True), (2, False)])
def test_splitting_of_one_dimensional_data(oracle, along_x, acc, crp):
    D = 4e-3
    base = _one_d(oracle, along_x, "NONE", acc, crp, D, 4)
    simple = _one_d(oracle, along_x, "SIMPLE", acc, crp, D, 4)
    halves = _one_d(oracle, along_x, "NONE", acc, crp, D / 2, 8)
    strang = _one_d(oracle, along_x, "STRANG", acc, crp, D, 4)
    # STRANG: X for dt / 2, Y for dt, X for dt / 2 -- two half steps of
    # x-only data, one whole step of y-only data
    expect, other = (halves, base) if along_x else (base, halves)
    for k in ("rho", "u", "v", "P", "e"):
        np.testing.assert_allclose(simple[k], base[k], rtol=1e-12, atol=1e-13, err_msg=f"SIMPLE {k}")
        np.testing.assert_allclose(strang[k], expect[k], rtol=1e-12, atol=1e-13, err_msg=f"STRANG {k}")
    assert not np.allclose(strang["rho"], other["rho"], rtol=1e-8, atol=0)
