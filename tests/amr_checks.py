"""Structural invariants of an AMR-2D mesh in the AmrCells layout (8 face
slots per cell, slot s of cell c at 8c + s), restating what the reference's
debug checks verify (EuMesh::check_refined, mesh/euler/eu_mesh.cpp:680-762;
check_flags, mesh/amr/balancing_restrictions.h:62-111)."""
import numpy as np

UNDEFINED, INNER, PERIODIC = -1, 0, 3
OPPOSITE = [1, 0, 3, 2]


def check_mesh(cells, slots, domain_area=None, max_level=None, tol=1e-12):
    """Raise AssertionError on the first violated invariant."""
    n = len(cells["cx"])
    bnd = slots["boundary"].reshape(n, 8)
    adj = slots["adj"].reshape(n, 8)
    fcx = slots["fcx"].reshape(n, 8)
    fcy = slots["fcy"].reshape(n, 8)
    area = slots["area"].reshape(n, 8)
    lvl = cells["level"]
    if domain_area is not None:
        assert abs(cells["volume"].sum() - domain_area) <= 1e-12 * domain_area * max(1, n ** 0.5)
    if max_level is not None:
        assert lvl.min() >= 0 and lvl.max() <= max_level
    for c in range(n):
        for side in range(4):
            s0, s1 = side, side + 4
            assert bnd[c, s0] != UNDEFINED, f"cell {c}: simple face {side} undefined"
            subs = [s0] if bnd[c, s1] == UNDEFINED else [s0, s1]
            for s in subs:
                b = bnd[c, s]
                j = adj[c, s]
                if b not in (INNER, PERIODIC):
                    assert j == c, f"cell {c} slot {s}: boundary slot must reference itself"
                    continue
                assert 0 <= j < n, f"cell {c} slot {s}: neighbour {j} out of range"
                assert abs(int(lvl[j]) - int(lvl[c])) <= 1, f"cell {c} slot {s}: 2:1 balance violated"
                if len(subs) == 2:
                    assert lvl[j] == lvl[c] + 1, f"cell {c} slot {s}: split face to a non-finer neighbour"
                # the neighbour has a slot on the opposite side pointing back
                opp = OPPOSITE[side]
                back = [t for t in (opp, opp + 4) if bnd[j, t] != UNDEFINED and adj[j, t] == c]
                assert back, f"cell {c} slot {s}: neighbour {j} does not point back"
                if b == INNER:
                    t = back[0] if len(back) == 1 else min(
                        back, key=lambda t: abs(fcx[j, t] - fcx[c, s]) + abs(fcy[j, t] - fcy[c, s]))
                    scale = max(area[c, s], area[j, t])
                    if lvl[j] == lvl[c] or len(subs) == 2 or bnd[j, opp + 4] != UNDEFINED:
                        assert abs(fcx[j, t] - fcx[c, s]) <= tol * 10 + 1e-9 * scale, (c, s, j, t)
                        assert abs(fcy[j, t] - fcy[c, s]) <= tol * 10 + 1e-9 * scale, (c, s, j, t)


def check_balanced_flags(cells, slots, flags, max_level):
    """After balancing: wanted levels of neighbours differ by <= 1, levels stay
    in [0, max_level], and siblings coarsen together."""
    n = len(cells["cx"])
    bnd = slots["boundary"].reshape(n, 8)
    adj = slots["adj"].reshape(n, 8)
    lvl = cells["level"]
    want = lvl + flags
    assert want.min() >= 0 and want.max() <= max_level
    for c in range(n):
        for s in range(8):
            if bnd[c, s] in (INNER, PERIODIC):
                assert abs(int(want[adj[c, s]]) - int(want[c])) <= 1, (c, s)
    z = cells["z_idx"]
    par = {}
    for c in np.flatnonzero(flags < 0):
        par.setdefault((int(cells["b_idx"][c]), int(lvl[c]), int(z[c]) // 4), []).append(c)
    for k, v in par.items():
        assert len(v) == 4, f"coarsening family {k} incomplete: {v}"
