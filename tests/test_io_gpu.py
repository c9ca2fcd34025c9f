"""VTU / PVD output and the checkpoint round trip (io/vtu_file.cpp,
io/pvd_file.h, EuMesh::backup):

* the VTU file parses as a VTK UnstructuredGrid with raw appended arrays:
  4 corner points per cell (or the polygon through the hanging-face
  midpoints), connectivity, offsets, cell types, and cell data equal to the
  context's downloads;
* the PVD collection lists every saved VTU with its time;
* a backup of an adapted mesh in the middle of a MUSCL run, restored into a
  fresh context, reproduces the mesh, state, isolation counters and solver
  parameters, and the restored run continues bit for bit like the original
  through further steps and adaptations (on a power-of-two base: an upload
  derives the LSM weights with glibc's pow like the reference, a device
  adaptation with the correctly rounded cube, which agree on dyadic
  geometry); a Rectangle restored from a backup runs the tiled step again.
"""
import json
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def adapted(zfv, seed=3, n=64):
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(n, n, bounds=(2, 1, 3, 3))
    ctx.configure(acc=2, flux="hllc", limiter="minmod", cfl=0.5, gamma=GAMMA)
    cells, _, _ = ctx.mesh_download()
    ctx.set_state(*problems.circle_blast(cells["cx"], cells["cy"], seed))
    for _ in range(3):
        ctx.set_flags(1, 0.1)
        ctx.refine(2, slope=True)
        ctx.step(3)
    return ctx


def read_vtu(path):
    raw = open(path, "rb").read()
    head, _, rest = raw.partition(b'<AppendedData encoding="raw">\n_')
    data = rest[: rest.rindex(b"\n  </AppendedData>")]
    root = ET.fromstring(head.decode() + "</VTKFile>")
    piece = root.find("UnstructuredGrid/Piece")
    dtypes = {"Float64": np.float64, "Int32": np.int32, "UInt8": np.uint8}
    out = {}
    for arr in piece.iter("DataArray"):
        off = int(arr.get("offset"))
        nbytes = int(np.frombuffer(data[off:off + 4], np.uint32)[0])
        a = np.frombuffer(data[off + 4:off + 4 + nbytes], dtypes[arr.get("type")])
        comps = int(arr.get("NumberOfComponents", "1"))
        out[arr.get("Name")] = a.reshape(-1, comps) if comps > 1 else a
    return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), out


@pytest.mark.parametrize("polygons", [False, True])
def test_vtu_file_matches_the_context(zfv, tmp_path, polygons):
    ctx = adapted(zfv)
    path = str(tmp_path / "mesh.vtu")
    ctx.vtu_write(path, zfv.VTU_ALL, polygons)
    npts, ncells, a = read_vtu(path)
    cells, slots, (vx, vy) = ctx.mesh_download()
    n = ctx.n
    assert ncells == n
    bnd = slots["boundary"].reshape(n, 8)
    split = bnd[:, 4:] >= 0
    if not polygons:
        assert npts == 4 * n and np.all(a["types"] == 9)
        corners = a["Points"].reshape(n, 4, 3)
        for q, k in enumerate((0, 2, 8, 6)):
            assert np.array_equal(corners[:, q, 0], vx.reshape(n, 9)[:, k])
            assert np.array_equal(corners[:, q, 1], vy.reshape(n, 9)[:, k])
    else:
        nv = 4 + split.sum(axis=1)
        assert np.array_equal(np.diff(np.concatenate([[0], a["offsets"]])), nv)
        assert np.all(a["types"] == np.where(nv == 4, 9, 7))
        assert split.any(), "the adapted mesh has hanging faces"
    assert np.array_equal(a["connectivity"], np.arange(npts))
    rho, u, v, p, e = ctx.state()
    assert np.array_equal(a["density"], rho) and np.array_equal(a["pressure"], p)
    assert np.array_equal(a["energy"], e)
    assert np.array_equal(a["velocity"][:, 0], u) and np.array_equal(a["velocity"][:, 1], v)
    assert np.all(a["velocity"][:, 2] == 0.0)
    assert np.array_equal(a["level"], cells["level"]) and np.array_equal(a["index"], np.arange(n))
    assert np.array_equal(a["wait"], ctx.wait())


def test_pvd_series(zfv, tmp_path):
    ctx = adapted(zfv, n=32)
    pvd = zfv.PvdFile("run", str(tmp_path))
    times = [0.0, 0.125, 0.5]
    for t in times:
        pvd.save(ctx, t, zfv.VTU_DENSITY | zfv.VTU_LEVEL)
        ctx.step(2)
    pvd.close()
    root = ET.parse(str(tmp_path / "run.pvd")).getroot()
    sets = root.findall("Collection/DataSet")
    assert [float(s.get("timestep")) for s in sets] == times
    for s in sets:
        f = tmp_path / s.get("file")
        assert f.exists()
        _, ncells, a = read_vtu(str(f))
        assert set(a) >= {"density", "level", "Points"} and "pressure" not in a


def test_backup_restore_continues_bit_for_bit(zfv, tmp_path):
    ctx = adapted(zfv)
    ctx.configure(acc=2, flux="hllc", limiter="minmod", cfl=0.45, gamma=GAMMA, max_dt=0.01)
    w = ctx.wait()
    w[::17] = 2
    ctx.set_wait(w)
    root = str(tmp_path / "ckpt")
    ctx.backup(root)
    desc = json.load(open(os.path.join(root, "backup.json")))
    assert desc["mesh"]["cells"]["size"] == ctx.n
    assert desc["mesh"]["cells"]["data"]["init"]["sizeof"] == 48
    res = zfv.Context()
    res.restore(root)
    assert res.params() == ctx.params()
    for a, b in zip(res.mesh_download()[0].values(), ctx.mesh_download()[0].values()):
        assert np.array_equal(a, b)
    for a, b in zip(res.state(), ctx.state()):
        assert np.array_equal(a, b)
    assert np.array_equal(res.wait(), ctx.wait())
    for cycle in range(3):
        for c in (ctx, res):
            c.step(4)
            c.set_flags(1, 0.1)
            c.refine(2, slope=True)
        assert res.dt() == ctx.dt()
        for a, b in zip(res.state(), ctx.state()):
            assert np.array_equal(a, b), f"cycle {cycle}"
        assert np.array_equal(res.mesh_download()[1]["adj"], ctx.mesh_download()[1]["adj"])


def test_restored_rectangle_runs_the_tiled_step(zfv, tmp_path):
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle(70, 40, bounds=(3, 3, 1, 2))
    ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.5, gamma=GAMMA)
    i, j = np.divmod(np.arange(70 * 40), 40)
    ctx.set_state(*problems.four_quadrant((i + 0.5) / 70, (j + 0.5) / 40, 5))
    ctx.step(3)
    ctx.backup(str(tmp_path / "rect"))
    res = zfv.Context()
    res.restore(str(tmp_path / "rect"))
    assert res.launch(1)[0] == zfv.TILE
    for c in (ctx, res):
        c.step(5)
    for a, b in zip(res.state(), ctx.state()):
        assert np.array_equal(a, b)


def test_io_errors(zfv, tmp_path):
    ctx = zfv.Context()
    with pytest.raises(zfv.ZfvError):
        ctx.vtu_write(str(tmp_path / "x.vtu"))   # no mesh
    with pytest.raises(zfv.ZfvError):
        ctx.restore(str(tmp_path / "missing"))


def test_axial_backup_restores_the_axial_run(zfv, tmp_path):
    """An axisymmetric mesh and solver: the backup holds volume_alt and
    area_alt (written when present, as AmrCells::backup does) and the
    axial flag; the restored run continues bit for bit."""
    from zephyr_amd import problems
    ctx = zfv.Context()
    ctx.mesh_rectangle_axial(40, 30, 0.0, 2.0, 0.0, 1.0, bounds=(2, 2, 2, 1))
    ctx.configure(acc=2, flux="hllc", limiter="MC", cfl=0.45, gamma=GAMMA, axial=True)
    i, j = np.divmod(np.arange(40 * 30), 30)
    ctx.set_state(*problems.four_quadrant((i + 0.5) / 40, (j + 0.5) / 30, 8))
    ctx.step(3)
    ctx.backup(str(tmp_path / "ax"))
    desc = json.load(open(tmp_path / "ax" / "backup.json"))
    assert desc["mesh"]["cells"]["axial"] is True and desc["solver"]["axial"] is True
    assert "volume_alt" in desc["mesh"]["cells"] and "area_alt" in desc["mesh"]["cells"]["faces"]
    res = zfv.Context()
    res.restore(str(tmp_path / "ax"))
    assert res.params()["axial"] == 1
    for a, b in zip(res.mesh_axial(), ctx.mesh_axial()):
        assert np.array_equal(a, b)
    for c in (ctx, res):
        c.step(4)
    for a, b in zip(res.state(), ctx.state()):
        assert np.array_equal(a, b)
