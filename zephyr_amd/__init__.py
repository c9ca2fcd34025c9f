"""zephyr_amd — MI355X-native finite-volume step of zephyr (SmFluid) on AMR-2D meshes.

Python plumbing over the C ABI of libzfv.so (include/zfv.h).  All compute
runs in the HIP kernels of the library; there is no CPU fallback: if the
library is missing or no GPU is present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import build as _build

LIB_PATH = _build.LIB_PATH

GODUNOV, RUSANOV, HLL, HLLC, HLLC_M, HLLC_LM, CIR1, CIR2 = 0, 1, 2, 3, 4, 5, 6, 7
FLUXES = {"godunov": GODUNOV, "rusanov": RUSANOV, "hll": HLL, "hllc": HLLC, "hllc_m": HLLC_M,
          "hllc_lm": HLLC_LM, "cir1": CIR1, "cir2": CIR2}
UNDEFINED, INNER, WALL, ZOE, PERIODIC, FUNCTION = -1, 0, 1, 2, 3, 4

ZFV_OK, ZFV_ERR_ARG, ZFV_ERR_HIP, ZFV_ERR_NUMERIC, ZFV_ERR_UNSUPPORTED, ZFV_ERR_STATE = range(6)

# Every symbol include/zfv.h declares.
EXPORTS = [
    "zfv_last_error", "zfv_version", "zfv_create", "zfv_destroy", "zfv_set_stream",
    "zfv_synchronize", "zfv_mesh_upload", "zfv_mesh_rectangle", "zfv_mesh_rectangle_axial",
    "zfv_mesh_axial_download", "zfv_mesh_upload_axial", "zfv_mesh_n_cells",
    "zfv_mesh_download", "zfv_smf_configure", "zfv_smf_set_max_dt", "zfv_state_upload",
    "zfv_state_download", "zfv_state_upload_device", "zfv_state_download_device",
    "zfv_wait_upload", "zfv_wait_download", "zfv_grad_download", "zfv_smf_step", "zfv_smf_dt",
    "zfv_smf_compute_grad", "zfv_smf_compute_dt", "zfv_conserved_totals", "zfv_conserved_totals_noalloc",
    "zfv_limiter_from_name", "zfv_profile_enable", "zfv_profile_read",
    "zfv_flags_upload", "zfv_flags_download", "zfv_set_flags", "zfv_balance_flags", "zfv_refine",
    "zfv_mesh_n_faces", "zfv_mesh_reserve", "zfv_set_launch", "zfv_get_launch", "zfv_autotune",
    "zfv_batch_rectangles", "zfv_batch_layout", "zfv_batch_dt", "zfv_graph_enable",
    "zfv_set_cell_order", "zfv_get_cell_order", "zfv_get_curve_order",
    "zfv_set_tile_schedule", "zfv_get_tile_schedule",
    "zfv_mm_set_limiter", "zfv_mm_set_accuracy", "zfv_mm_set_method", "zfv_mm_set_crp_mode",
    "zfv_mm_set_splitting", "zfv_mm_set_overlap", "zfv_mm_set_flags", "zfv_mm_refine", "zfv_mm_grad",
    "zfv_adjoint_forward", "zfv_adjoint_backward",
    "zfv_mm_configure", "zfv_mm_set_crp", "zfv_mm_set_max_dt", "zfv_mm_set_state", "zfv_mm_state", "zfv_mm_step", "zfv_mm_dt", "zfv_mm_crp_faces",
    "zfv_mm_interface", "zfv_mm_domain",
    "zfv_transfer_configure", "zfv_transfer_set_velocity", "zfv_transfer_set_state", "zfv_transfer_state",
    "zfv_transfer_slopes", "zfv_transfer_set_dt", "zfv_transfer_get_dt", "zfv_transfer_compute_dt",
    "zfv_transfer_update", "zfv_transfer_set_flags",
    "zfv_smf_get_params", "zfv_vtu_write", "zfv_pvd_open", "zfv_pvd_save", "zfv_pvd_close",
    "zfv_backup", "zfv_restore",
    "zfv3_create", "zfv3_destroy", "zfv3_mesh_cuboid", "zfv3_mesh_n_cells", "zfv3_mesh_download",
    "zfv3_smf_configure", "zfv3_smf_set_max_dt", "zfv3_state_upload", "zfv3_state_download",
    "zfv3_wait_upload", "zfv3_wait_download", "zfv3_grad_download", "zfv3_smf_step", "zfv3_smf_dt",
    "zfv3_smf_compute_dt", "zfv3_conserved_totals", "zfv3_synchronize", "zfv3_profile_enable",
    "zfv3_profile_read", "zfv3_set_geometry_tables", "zfv3_get_geometry_tables",
]
VTU_DENSITY, VTU_VELOCITY, VTU_PRESSURE, VTU_ENERGY, VTU_WAIT, VTU_LEVEL, VTU_FLAG, VTU_INDEX = (
    1, 2, 4, 8, 16, 32, 64, 128)
VTU_ALL = 255
CELL, QUAD, BATCH, TILE = 0, 1, 2, 3   # step-kernel layouts (TILE: the whole MUSCL step, stage 1)


class ZfvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"zfv error {code}: {msg}")
        self.code = code


class MeshSoA(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int32),
        ("cx", C.c_void_p), ("cy", C.c_void_p), ("volume", C.c_void_p),
        ("vx", C.c_void_p), ("vy", C.c_void_p),
        ("level", C.c_void_p), ("flag", C.c_void_p), ("b_idx", C.c_void_p), ("z_idx", C.c_void_p),
        ("boundary", C.c_void_p), ("adj", C.c_void_p),
        ("nx", C.c_void_p), ("ny", C.c_void_p), ("fcx", C.c_void_p), ("fcy", C.c_void_p),
        ("area", C.c_void_p),
    ]


class SmfParams(C.Structure):
    _fields_ = [
        ("acc", C.c_int32), ("flux", C.c_int32), ("limiter", C.c_int32), ("axial", C.c_int32),
        ("cfl", C.c_double), ("max_dt", C.c_double),
        ("gamma", C.c_double), ("P0", C.c_double), ("e0", C.c_double),
    ]


_lib = None


def lib():
    """Load libzfv.so (build it first if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            _build.build()
        L = C.CDLL(LIB_PATH)
        vp, i32, d = C.c_void_p, C.c_int32, C.c_double
        L.zfv_last_error.restype = C.c_char_p
        L.zfv_version.restype = C.c_char_p
        L.zfv_limiter_from_name.restype = C.c_int
        L.zfv_limiter_from_name.argtypes = [C.c_char_p]
        for name in EXPORTS:
            f = getattr(L, name)
            if name not in ("zfv_last_error", "zfv_version", "zfv_limiter_from_name"):
                f.restype = C.c_int
        L.zfv_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.zfv_mesh_rectangle.argtypes = [vp, C.c_int, C.c_int, d, d, d, d, vp]
        L.zfv_mesh_rectangle_axial.argtypes = [vp, C.c_int, C.c_int, d, d, d, d, vp]
        L.zfv_mesh_axial_download.argtypes = [vp, vp, vp]
        L.zfv_smf_set_max_dt.argtypes = [vp, d]
        L.zfv_smf_step.argtypes = [vp, C.c_int]
        L.zfv_profile_enable.argtypes = [vp, C.c_int]
        L.zfv_set_flags.argtypes = [vp, C.c_int, d]
        L.zfv_balance_flags.argtypes = [vp, C.c_int]
        L.zfv_refine.argtypes = [vp, C.c_int, C.c_int]
        L.zfv_mesh_reserve.argtypes = [vp, C.c_int]
        L.zfv_vtu_write.argtypes = [vp, C.c_char_p, C.c_uint, C.c_int]
        L.zfv_pvd_open.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp)]
        L.zfv_pvd_save.argtypes = [vp, vp, d, C.c_uint, C.c_int]
        L.zfv_pvd_close.argtypes = [vp]
        L.zfv_backup.argtypes = [vp, C.c_char_p]
        L.zfv_restore.argtypes = [vp, C.c_char_p]
        L.zfv_smf_get_params.argtypes = [vp, vp]
        L.zfv_set_launch.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.zfv_autotune.argtypes = [vp, C.c_int, vp]
        L.zfv_batch_rectangles.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.zfv_graph_enable.argtypes = [vp, C.c_int]
        L.zfv_set_cell_order.argtypes = [vp, C.c_int]
        L.zfv_get_cell_order.argtypes = [vp, vp, vp]
        L.zfv_get_curve_order.argtypes = [vp, vp]
        L.zfv_set_tile_schedule.argtypes = [vp, C.c_int]
        L.zfv_get_tile_schedule.argtypes = [vp, vp]
        L.zfv_mm_configure.argtypes = [vp, C.c_int, vp, C.c_double]
        L.zfv_mm_set_max_dt.argtypes = [vp, C.c_double]
        L.zfv_mm_set_crp.argtypes = [vp, C.c_int]
        L.zfv_mm_set_state.argtypes = [vp] * 6
        L.zfv_mm_state.argtypes = [vp] * 9
        L.zfv_mm_step.argtypes = [vp, C.c_int]
        L.zfv_mm_dt.argtypes = [vp, vp]
        L.zfv_mm_interface.argtypes = [vp] * 4
        L.zfv_mm_domain.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
        L.zfv_transfer_configure.argtypes = [vp, C.c_int, d]
        L.zfv_transfer_set_velocity.argtypes = [vp] * 5
        L.zfv_transfer_set_state.argtypes = [vp, vp]
        L.zfv_transfer_state.argtypes = [vp] * 4
        L.zfv_transfer_slopes.argtypes = [vp] * 3
        L.zfv_transfer_set_dt.argtypes = [vp, d]
        L.zfv_transfer_get_dt.argtypes = [vp, vp]
        L.zfv_transfer_compute_dt.argtypes = [vp, vp]
        L.zfv_transfer_update.argtypes = [vp, C.c_int, C.c_int]
        L.zfv_transfer_set_flags.argtypes = [vp]
        L.zfv_mm_set_limiter.argtypes = [vp, C.c_int]
        L.zfv_mm_set_accuracy.argtypes = [vp, C.c_int]
        L.zfv_mm_set_method.argtypes = [vp, C.c_int]
        L.zfv_mm_set_crp_mode.argtypes = [vp, C.c_int]
        L.zfv_mm_set_splitting.argtypes = [vp, C.c_int]
        L.zfv_mm_set_overlap.argtypes = [vp, C.c_int, C.c_int]
        L.zfv_mm_set_flags.argtypes = [vp]
        L.zfv_mm_refine.argtypes = [vp, C.c_int]
        L.zfv_mm_grad.argtypes = [vp, vp, vp]
        L.zfv_adjoint_forward.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        L.zfv_adjoint_backward.argtypes = [vp, vp, vp, vp, vp, vp]
        L.zfv3_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.zfv3_mesh_cuboid.argtypes = [vp, C.c_int, C.c_int, C.c_int, d, d, d, d, d, d, vp]
        L.zfv3_mesh_download.argtypes = [vp] * 16
        L.zfv3_smf_set_max_dt.argtypes = [vp, d]
        L.zfv3_state_upload.argtypes = [vp] * 7
        L.zfv3_state_download.argtypes = [vp] * 7
        L.zfv3_smf_step.argtypes = [vp, C.c_int]
        L.zfv3_profile_enable.argtypes = [vp, C.c_int]
        L.zfv3_set_geometry_tables.argtypes = [vp, C.c_int]
        _lib = L
    return _lib


def _check(code):
    if code != ZFV_OK:
        raise ZfvError(code, lib().zfv_last_error().decode())


def limiter_id(name):
    if isinstance(name, int):
        return name
    v = lib().zfv_limiter_from_name(name.encode())
    if v < 0:
        raise ValueError(f"Unknown limiter '{name}'")
    return v


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class PvdFile:
    """PvdFile (io/pvd_file.h:12-75): a series of VTU files and their PVD collection."""

    def __init__(self, filename, directory="output"):
        self.h = C.c_void_p()
        _check(lib().zfv_pvd_open(os.fsencode(filename), os.fsencode(directory), C.byref(self.h)))

    def save(self, ctx, time, fields=VTU_ALL, polygons=False):
        _check(lib().zfv_pvd_save(self.h, ctx.h, float(time), fields, 1 if polygons else 0))

    def close(self):
        if self.h:
            lib().zfv_pvd_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device mesh + SmFluid solver state (a zfv_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().zfv_create(device, C.byref(h)))
        self.h = h
        self.n = 0
        self._keep = []

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().zfv_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- mesh
    def mesh_rectangle(self, nx, ny, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, bounds=(ZOE,) * 4):
        b = np.ascontiguousarray(bounds, dtype=np.int32)
        _check(lib().zfv_mesh_rectangle(self.h, nx, ny, xmin, xmax, ymin, ymax, _ptr(b)))
        self.n = nx * ny

    def mesh_rectangle_axial(self, nx, ny, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, bounds=(ZOE,) * 4):
        """Rectangle with set_axial(true): y is the radius."""
        b = np.ascontiguousarray(bounds, dtype=np.int32)
        _check(lib().zfv_mesh_rectangle_axial(self.h, nx, ny, xmin, xmax, ymin, ymax, _ptr(b)))
        self.n = nx * ny

    def mesh_axial(self):
        """(volume_alt (n), area_alt (8n)) of an axisymmetric mesh."""
        v, a = np.zeros(self.n), np.zeros(8 * self.n)
        _check(lib().zfv_mesh_axial_download(self.h, _ptr(v), _ptr(a)))
        return v, a

    def mesh_upload(self, cells, slots, verts):
        n = len(cells["cx"])
        arrs = dict(
            cx=np.ascontiguousarray(cells["cx"], np.float64), cy=np.ascontiguousarray(cells["cy"], np.float64),
            volume=np.ascontiguousarray(cells["volume"], np.float64),
            vx=np.ascontiguousarray(verts[0], np.float64), vy=np.ascontiguousarray(verts[1], np.float64),
            level=np.ascontiguousarray(cells.get("level", np.zeros(n)), np.int32),
            flag=np.ascontiguousarray(cells.get("flag", np.zeros(n)), np.int32),
            b_idx=np.ascontiguousarray(cells.get("b_idx", np.arange(n)), np.int32),
            z_idx=np.ascontiguousarray(cells.get("z_idx", np.zeros(n)), np.int32),
            boundary=np.ascontiguousarray(slots["boundary"], np.int32),
            adj=np.ascontiguousarray(slots["adj"], np.int32),
            nx=np.ascontiguousarray(slots["nx"], np.float64), ny=np.ascontiguousarray(slots["ny"], np.float64),
            fcx=np.ascontiguousarray(slots["fcx"], np.float64), fcy=np.ascontiguousarray(slots["fcy"], np.float64),
            area=np.ascontiguousarray(slots["area"], np.float64),
        )
        m = MeshSoA(n_cells=n, **{k: v.ctypes.data for k, v in arrs.items()})
        _check(lib().zfv_mesh_upload(self.h, C.byref(m)))
        self.n = n

    def n_faces(self):
        v = C.c_int64()
        _check(lib().zfv_mesh_n_faces(self.h, C.byref(v)))
        return v.value

    def vtu_write(self, filename, fields=VTU_ALL, polygons=False):
        """VtuFile::save: the mesh and cell fields as a VTU file."""
        _check(lib().zfv_vtu_write(self.h, os.fsencode(filename), fields, 1 if polygons else 0))

    def backup(self, root):
        """EuMesh::backup of the mesh and the solver state into directory root."""
        _check(lib().zfv_backup(self.h, os.fsencode(root)))

    def restore(self, root):
        """Mesh, solver parameters and state from a zfv_backup directory."""
        _check(lib().zfv_restore(self.h, os.fsencode(root)))
        n = C.c_int32()
        _check(lib().zfv_mesh_n_cells(self.h, C.byref(n)))
        self.n = n.value

    def params(self):
        """The solver's parameters (acc, flux, limiter, axial, cfl, max_dt, gamma, P0, e0)."""
        p = SmfParams()
        _check(lib().zfv_smf_get_params(self.h, C.byref(p)))
        return dict(acc=p.acc, flux=p.flux, limiter=p.limiter, axial=p.axial, cfl=p.cfl, max_dt=p.max_dt,
                    gamma=p.gamma, P0=p.P0, e0=p.e0)

    def reserve(self, n_cells):
        """Device capacity for meshes of up to n_cells cells (AmrCells::reserve_amr)."""
        _check(lib().zfv_mesh_reserve(self.h, int(n_cells)))

    def mesh_download(self):
        n = C.c_int32()
        _check(lib().zfv_mesh_n_cells(self.h, C.byref(n)))
        n = n.value
        cells = dict(cx=np.zeros(n), cy=np.zeros(n), volume=np.zeros(n),
                     level=np.zeros(n, np.int32), flag=np.zeros(n, np.int32),
                     b_idx=np.zeros(n, np.int32), z_idx=np.zeros(n, np.int32))
        slots = dict(boundary=np.zeros(8 * n, np.int32), adj=np.zeros(8 * n, np.int32),
                     nx=np.zeros(8 * n), ny=np.zeros(8 * n), fcx=np.zeros(8 * n), fcy=np.zeros(8 * n),
                     area=np.zeros(8 * n))
        vx, vy = np.zeros(9 * n), np.zeros(9 * n)
        m = MeshSoA(n_cells=n, vx=vx.ctypes.data, vy=vy.ctypes.data,
                    **{k: v.ctypes.data for k, v in list(cells.items()) + list(slots.items())})
        _check(lib().zfv_mesh_download(self.h, C.byref(m)))
        return cells, slots, (vx, vy)

    # ---- solver
    def configure(self, acc=1, flux=HLLC, limiter="MC", cfl=0.5, gamma=1.4, P0=0.0, e0=0.0, max_dt=None,
                  axial=False):
        """SmFluid setters; axial = set_axial (needs mesh_rectangle_axial)."""
        if isinstance(flux, str):
            flux = FLUXES[flux.lower()]
        p = SmfParams(acc=acc, flux=flux, limiter=limiter_id(limiter), axial=1 if axial else 0, cfl=cfl,
                      max_dt=np.finfo(np.float64).max if max_dt is None else max_dt,
                      gamma=gamma, P0=P0, e0=e0)
        _check(lib().zfv_smf_configure(self.h, C.byref(p)))

    def set_max_dt(self, v):
        _check(lib().zfv_smf_set_max_dt(self.h, v))

    def set_state(self, rho, u, v, p, e):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, p, e)]
        _check(lib().zfv_state_upload(self.h, *[_ptr(a) for a in arrs]))

    def set_state_device(self, ptrs):
        _check(lib().zfv_state_upload_device(self.h, *[C.c_void_p(p) for p in ptrs]))

    def state(self):
        arrs = [np.zeros(self.n) for _ in range(5)]
        _check(lib().zfv_state_download(self.h, *[_ptr(a) for a in arrs]))
        return arrs

    def set_wait(self, w):
        w = np.ascontiguousarray(w, dtype=np.int32)
        _check(lib().zfv_wait_upload(self.h, _ptr(w)))

    def wait(self):
        w = np.zeros(self.n, np.int32)
        _check(lib().zfv_wait_download(self.h, _ptr(w)))
        return w

    def grad(self):
        dx, dy = np.zeros(4 * self.n), np.zeros(4 * self.n)
        _check(lib().zfv_grad_download(self.h, _ptr(dx), _ptr(dy)))
        return dx.reshape(4, self.n), dy.reshape(4, self.n)

    # ---- adaptation
    def set_flags_array(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        _check(lib().zfv_flags_upload(self.h, _ptr(f)))

    def flags(self):
        n = C.c_int32()
        _check(lib().zfv_mesh_n_cells(self.h, C.byref(n)))
        f = np.zeros(n.value, np.int32)
        _check(lib().zfv_flags_download(self.h, _ptr(f)))
        return f

    def set_flags(self, criterion=0, threshold=0.1):
        _check(lib().zfv_set_flags(self.h, criterion, threshold))

    def balance_flags(self, max_level):
        _check(lib().zfv_balance_flags(self.h, max_level))

    def refine(self, max_level, slope=True):
        _check(lib().zfv_refine(self.h, max_level, 1 if slope else 0))
        n = C.c_int32()
        _check(lib().zfv_mesh_n_cells(self.h, C.byref(n)))
        self.n = n.value

    def step(self, n=1):
        _check(lib().zfv_smf_step(self.h, n))

    def dt(self):
        v = C.c_double()
        _check(lib().zfv_smf_dt(self.h, C.byref(v)))
        return v.value

    def compute_grad(self):
        _check(lib().zfv_smf_compute_grad(self.h))

    def compute_dt(self):
        v = C.c_double()
        _check(lib().zfv_smf_compute_dt(self.h, C.byref(v)))
        return v.value

    def totals(self):
        out = np.zeros(4)
        _check(lib().zfv_conserved_totals(self.h, _ptr(out)))
        return out

    def totals_noalloc(self):
        """zfv_conserved_totals_noalloc: the same totals without device allocation."""
        out = np.zeros(4)
        _check(lib().zfv_conserved_totals_noalloc(self.h, _ptr(out)))
        return out

    def synchronize(self):
        _check(lib().zfv_synchronize(self.h))

    def set_stream(self, stream_ptr):
        _check(lib().zfv_set_stream(self.h, C.c_void_p(stream_ptr)))

    def set_launch(self, stage, layout, block=0):
        _check(lib().zfv_set_launch(self.h, stage, layout, block))

    def launch(self, stage):
        """(layout, workgroup size) of a stage."""
        v, b = C.c_int(), C.c_int()
        _check(lib().zfv_get_launch(self.h, stage, C.byref(v), C.byref(b)))
        return v.value, b.value

    def autotune(self, reps=5):
        """ms per launch, [stage][layout 0-3][block 128/256/512, or the tile
        shape 512/768/1024 for layout 3]."""
        ms = np.zeros(36)
        _check(lib().zfv_autotune(self.h, reps, _ptr(ms)))
        return ms.reshape(3, 4, 3)

    def graph_enable(self, on=True):
        _check(lib().zfv_graph_enable(self.h, 1 if on else 0))

    def set_tile_schedule(self, persistent):
        """Tiled step by persistent workgroups (True) or one workgroup per
        tile (False); results are identical."""
        _check(lib().zfv_set_tile_schedule(self.h, 1 if persistent else 0))

    def tile_schedule(self):
        p = C.c_int(0)
        _check(lib().zfv_get_tile_schedule(self.h, C.byref(p)))
        return bool(p.value)

    def set_cell_order(self, curve=True, rebuild=False):
        """Curve order for the step kernels of refined meshes (default;
        carried across adaptations, or with rebuild=True computed afresh
        after each) or the mesh's own order throughout; results are
        identical."""
        _check(lib().zfv_set_cell_order(self.h, (2 if rebuild else 1) if curve else 0))

    def curve_order(self):
        """The steps' cell order: mesh index of each cell along the curve."""
        out = np.empty(self.n, np.int32)
        _check(lib().zfv_get_curve_order(self.h, out.ctypes.data))
        return out

    def cell_order(self):
        """(mode, active): mode 1 = curve order for steps; active = the
        state is currently held in curve order."""
        m, a = C.c_int(0), C.c_int(0)
        _check(lib().zfv_get_cell_order(self.h, C.byref(m), C.byref(a)))
        return m.value, a.value

    # ---- multi-material solver
    def mm_configure(self, materials, cfl=0.7, crp=True):
        """materials: sequence of (gamma, P0, e0, Cv, T0, rho_ref); crp:
        composite Riemann problems at mixed faces (the MmFluid default)."""
        mats = np.ascontiguousarray(np.asarray(materials, dtype=np.float64).reshape(-1, 6))
        self._mm_nmat = mats.shape[0]
        _check(lib().zfv_mm_configure(self.h, mats.shape[0], _ptr(mats), float(cfl)))
        _check(lib().zfv_mm_set_crp(self.h, 1 if crp else 0))

    def mm_set_max_dt(self, v):
        _check(lib().zfv_mm_set_max_dt(self.h, float(v)))

    def mm_set_state(self, rho, u, v, p, beta):
        """beta: (n_mat, n) mass fractions."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, p)]
        b = np.ascontiguousarray(np.asarray(beta, dtype=np.float64).reshape(self._mm_nmat, -1))
        _check(lib().zfv_mm_set_state(self.h, *[_ptr(a) for a in arrs], _ptr(b)))

    def mm_state(self):
        """dict of rho, u, v, P, e, T (n) and beta, rhos (n_mat, n)."""
        n, nm = self.n, self._mm_nmat
        out = {k: np.zeros(n) for k in ("rho", "u", "v", "P", "e", "T")}
        out["beta"] = np.zeros((nm, n))
        out["rhos"] = np.zeros((nm, n))
        _check(lib().zfv_mm_state(self.h, *[_ptr(out[k]) for k in ("rho", "u", "v", "P", "e", "T", "beta", "rhos")]))
        return out

    def mm_step(self, n=1):
        _check(lib().zfv_mm_step(self.h, n))

    def mm_dt(self):
        d = C.c_double()
        _check(lib().zfv_mm_dt(self.h, C.byref(d)))
        return d.value

    def mm_crp_faces(self):
        """Cell sides that took a composite Riemann flux in the last step."""
        k = C.c_int64()
        _check(lib().zfv_mm_crp_faces(self.h, C.byref(k)))
        return k.value

    def mm_interface(self):
        """PLIC normals and section positions, (n_mat, n) each."""
        nm, n = self._mm_nmat, self.n
        nx, ny, p = np.zeros((nm, n)), np.zeros((nm, n)), np.zeros((nm, n))
        _check(lib().zfv_mm_interface(self.h, _ptr(nx), _ptr(ny), _ptr(p)))
        return nx, ny, p

    def mm_set_limiter(self, limiter):
        lim = limiter if isinstance(limiter, int) else lib().zfv_limiter_from_name(limiter.encode())
        _check(lib().zfv_mm_set_limiter(self.h, lim))

    def mm_set_accuracy(self, acc):
        """1: the first-order update (default); 2: MmFluid's second order
        (compute_grad, fluxes_stage1 / fluxes_stage2)."""
        _check(lib().zfv_mm_set_accuracy(self.h, int(acc)))

    MM_FLUXES = {"Godunov": 0, "Rusanov": 1, "HLL": 2, "HLLC": 3, "CRP": 8}   # Fluxes (num_flux.h:11-21)
    CRP_MODES = ("NONE", "V3", "V5", "VS", "PLIC", "MUSCL")                  # CrpMode (mm_fluid.h:36-42)
    SPLITS = ("NONE", "SIMPLE", "STRANG")                                    # DirSplit (mm_fluid.h:29-33)

    # ---- Transfer
    TRANSFER_METHODS = ["CRP_V3", "CRP_V5", "CRP_SE", "CRP_N1", "CRP_N2", "VOF", "VOF_CRP", "MUSCLd", "MUSCLn",
                        "MUSCLd_CRP", "MUSCLn_CRP", "MUSCL_MC", "MUSCL_MC_CRP", "WENO", "WENO_CRP"]

    def transfer_configure(self, method="MUSCL_MC", cfl=0.5):
        k = method if isinstance(method, int) else self.TRANSFER_METHODS.index(method)
        _check(lib().zfv_transfer_configure(self.h, k, cfl))

    def transfer_set_velocity(self, fvx, fvy, cvx, cvy):
        """Velocity samples at the face centres (8n) and cell centres (n)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (fvx, fvy, cvx, cvy)]
        _check(lib().zfv_transfer_set_velocity(self.h, *[_ptr(a) for a in arrs]))

    def transfer_set_state(self, u):
        _check(lib().zfv_transfer_set_state(self.h, _ptr(np.ascontiguousarray(u, dtype=np.float64))))

    def transfer_state(self):
        """(u, nx, ny): the fraction and the interface normals."""
        u, nx, ny = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        _check(lib().zfv_transfer_state(self.h, _ptr(u), _ptr(nx), _ptr(ny)))
        return u, nx, ny

    def transfer_slopes(self):
        dx, dy = np.zeros(self.n), np.zeros(self.n)
        _check(lib().zfv_transfer_slopes(self.h, _ptr(dx), _ptr(dy)))
        return dx, dy

    def transfer_set_dt(self, dt):
        _check(lib().zfv_transfer_set_dt(self.h, dt))

    def transfer_dt(self):
        v = C.c_double()
        _check(lib().zfv_transfer_get_dt(self.h, C.byref(v)))
        return v.value

    def transfer_compute_dt(self):
        v = C.c_double()
        _check(lib().zfv_transfer_compute_dt(self.h, C.byref(v)))
        return v.value

    def transfer_update(self, direction="ANY", n_steps=1):
        d = direction if isinstance(direction, int) else ["ANY", "X", "Y"].index(direction)
        _check(lib().zfv_transfer_update(self.h, d, n_steps))

    def transfer_set_flags(self):
        _check(lib().zfv_transfer_set_flags(self.h))

    def mm_domain(self, idx):
        """MmFluid::domain: (offsets, xs, ys) of the polygons material idx
        occupies, after interface_recovery of the current state."""
        npoly, nvert = C.c_int64(), C.c_int64()
        _check(lib().zfv_mm_domain(self.h, idx, C.byref(npoly), C.byref(nvert), None, None, None))
        off = np.zeros(npoly.value + 1, np.int32)
        xs, ys = np.zeros(nvert.value), np.zeros(nvert.value)
        _check(lib().zfv_mm_domain(self.h, idx, None, None, _ptr(off), _ptr(xs), _ptr(ys)))
        return off, xs, ys

    def mm_set_method(self, flux):
        """MmFluid::set_method: "CRP" (PLIC + HLLC) or a plain mmf flux
        ("HLLC", "HLL", "Rusanov", "Godunov") without composite fluxes; or a
        Fluxes id."""
        _check(lib().zfv_mm_set_method(self.h, flux if isinstance(flux, int) else self.MM_FLUXES[flux]))

    def mm_set_crp_mode(self, mode):
        """MmFluid::set_crp_mode by name ("V3", "V5", "VS", "PLIC"; "NONE" is
        ignored, as in the reference)."""
        _check(lib().zfv_mm_set_crp_mode(self.h, self.CRP_MODES.index(mode)))

    def mm_set_overlap(self, mode=0, crp_groups=0):
        """0 (default): one stream; 1: the interface on a second stream; 2:
        also the composite fluxes, beside the update of the other cells;
        crp_groups 0 = the default grid of the composite-flux kernel."""
        _check(lib().zfv_mm_set_overlap(self.h, int(mode), crp_groups))

    def mm_set_splitting(self, split):
        """MmFluid::set_splitting: "NONE", "SIMPLE" or "STRANG"."""
        _check(lib().zfv_mm_set_splitting(self.h, self.SPLITS.index(split)))

    def mm_set_flags(self):
        """MmFluid::set_flags into the mesh's flags (and the gradients kept for mm_refine)."""
        _check(lib().zfv_mm_set_flags(self.h))

    def mm_refine(self, max_level):
        """EuMesh::refine with MmFluid's distributor (after mm_set_flags)."""
        _check(lib().zfv_mm_refine(self.h, max_level))
        n = C.c_int32()
        _check(lib().zfv_mesh_n_cells(self.h, C.byref(n)))
        self.n = n.value

    def mm_grad(self):
        """Gradients of the last compute_grad (mm_set_flags or a second-order
        step): (gx, gy), rows rho, u, v, e, beta[n_mat], P, T, rhos[n_mat]."""
        K, n = 6 + 2 * self._mm_nmat, self.n
        gx, gy = np.zeros((K, n)), np.zeros((K, n))
        _check(lib().zfv_mm_grad(self.h, _ptr(gx), _ptr(gy)))
        return gx, gy

    # ---- discrete adjoint

    def adjoint_forward(self, n_steps, every, window):
        """n_steps steps with a checkpoint every `every`; returns J = int_window p^2 dA."""
        win = np.ascontiguousarray(window, np.float64)
        J = C.c_double()
        _check(lib().zfv_adjoint_forward(self.h, n_steps, every, _ptr(win), C.byref(J)))
        return J.value

    def adjoint_backward(self):
        """dJ/d{rho, u, v, p, e} of the state the last adjoint_forward started from."""
        arrs = [np.zeros(self.n) for _ in range(5)]
        _check(lib().zfv_adjoint_backward(self.h, *[_ptr(a) for a in arrs]))
        return arrs

    # ---- batches
    def batch_rectangles(self, sizes, boxes=None, bounds=None):
        """Independent Rectangle problems sizes = [(nx, ny), ...] in one context."""
        nx = np.ascontiguousarray([a for a, _ in sizes], np.int32)
        ny = np.ascontiguousarray([b for _, b in sizes], np.int32)
        bx = None if boxes is None else np.ascontiguousarray(boxes, np.float64).reshape(-1)
        bd = None if bounds is None else np.ascontiguousarray(bounds, np.int32).reshape(-1)
        _check(lib().zfv_batch_rectangles(self.h, len(sizes), _ptr(nx), _ptr(ny), _ptr(bx), _ptr(bd)))
        n = C.c_int32()
        _check(lib().zfv_mesh_n_cells(self.h, C.byref(n)))
        self.n = n.value

    def batch_layout(self):
        """(begin[np + 1], count[np]) cell ranges of the problems."""
        npb = C.c_int32()
        _check(lib().zfv_batch_layout(self.h, C.byref(npb), None, None))
        begin = np.zeros(npb.value + 1, np.int32)
        count = np.zeros(npb.value, np.int32)
        _check(lib().zfv_batch_layout(self.h, None, _ptr(begin), _ptr(count)))
        return begin, count

    def batch_dt(self):
        begin, _ = self.batch_layout()
        dt = np.zeros(len(begin) - 1)
        _check(lib().zfv_batch_dt(self.h, _ptr(dt)))
        return dt

    def profile_enable(self, on=True):
        _check(lib().zfv_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        cap = 16
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        cnt = (C.c_int32 * cap)()
        count = C.c_int32(cap)
        _check(lib().zfv_profile_read(self.h, names, ms, cnt, C.byref(count)))
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(count.value)}


class Context3:
    """A three-dimensional Cuboid mesh + SmFluid state on the device (a zfv3_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().zfv3_create(device, C.byref(h)))
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().zfv3_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mesh_cuboid(self, nx, ny, nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                    bounds=(ZOE,) * 6):
        """Cuboid (geom/generator/cuboid.cpp): bounds left, right, bottom, top, back, front."""
        b = np.ascontiguousarray(bounds, dtype=np.int32)
        _check(lib().zfv3_mesh_cuboid(self.h, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, _ptr(b)))
        self.n = nx * ny * nz

    def mesh(self):
        """(cells, slots): cells cx, cy, cz, vol, h (n); slots boundary, adj, nx, ny, nz, fcx, fcy,
        fcz, area, w (24 per cell, cell-major)."""
        n = self.n
        cells = {k: np.zeros(n) for k in ("cx", "cy", "cz", "vol", "h")}
        slots = {"boundary": np.zeros(24 * n, np.int32), "adj": np.zeros(24 * n, np.int32)}
        slots.update({k: np.zeros(24 * n) for k in ("nx", "ny", "nz", "fcx", "fcy", "fcz", "area", "w")})
        args = [cells[k] for k in ("cx", "cy", "cz", "vol", "h")] + [
            slots[k] for k in ("boundary", "adj", "nx", "ny", "nz", "fcx", "fcy", "fcz", "area", "w")]
        _check(lib().zfv3_mesh_download(self.h, *[_ptr(a) for a in args]))
        return cells, slots

    def configure(self, acc=1, flux=HLLC, limiter="MC", cfl=0.5, gamma=1.4, P0=0.0, e0=0.0, max_dt=None):
        if isinstance(flux, str):
            flux = FLUXES[flux.lower()]
        p = SmfParams(acc=acc, flux=flux, limiter=limiter_id(limiter), axial=0, cfl=cfl,
                      max_dt=np.finfo(np.float64).max if max_dt is None else max_dt,
                      gamma=gamma, P0=P0, e0=e0)
        _check(lib().zfv3_smf_configure(self.h, C.byref(p)))

    def set_max_dt(self, v):
        _check(lib().zfv3_smf_set_max_dt(self.h, v))

    def set_state(self, rho, u, v, w, p, e=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, w, p)]
        ea = None if e is None else np.ascontiguousarray(e, dtype=np.float64)
        _check(lib().zfv3_state_upload(self.h, *[_ptr(a) for a in arrs], _ptr(ea)))

    def state(self):
        """rho, u, v, w, p, e."""
        arrs = [np.zeros(self.n) for _ in range(6)]
        _check(lib().zfv3_state_download(self.h, *[_ptr(a) for a in arrs]))
        return arrs

    def set_wait(self, w):
        w = np.ascontiguousarray(w, dtype=np.int32)
        _check(lib().zfv3_wait_upload(self.h, _ptr(w)))

    def wait(self):
        w = np.zeros(self.n, np.int32)
        _check(lib().zfv3_wait_download(self.h, _ptr(w)))
        return w

    def grad(self):
        """[3][5][n]: d/dx, d/dy, d/dz of rho, u, v, w, p."""
        g = np.zeros(15 * self.n)
        _check(lib().zfv3_grad_download(self.h, _ptr(g)))
        return g.reshape(3, 5, self.n)

    def step(self, n=1):
        _check(lib().zfv3_smf_step(self.h, n))

    def dt(self):
        v = C.c_double()
        _check(lib().zfv3_smf_dt(self.h, C.byref(v)))
        return v.value

    def compute_dt(self):
        v = C.c_double()
        _check(lib().zfv3_smf_compute_dt(self.h, C.byref(v)))
        return v.value

    def totals(self):
        """mass, momentum x, y, z, energy."""
        out = np.zeros(5)
        _check(lib().zfv3_conserved_totals(self.h, _ptr(out)))
        return out

    def synchronize(self):
        _check(lib().zfv3_synchronize(self.h))

    def profile_enable(self, on=True):
        _check(lib().zfv3_profile_enable(self.h, 1 if on else 0))

    def set_geometry_tables(self, on=True):
        """Use the Cuboid's 1D geometry tables for meshes built afterwards."""
        _check(lib().zfv3_set_geometry_tables(self.h, 1 if on else 0))

    def geometry_tables(self):
        v = C.c_int()
        _check(lib().zfv3_get_geometry_tables(self.h, C.byref(v)))
        return bool(v.value)

    def profile_read(self):
        """{kernel: (total ms, launches)} since enabled or last read."""
        cap = 8
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        cnt = (C.c_int32 * cap)()
        count = C.c_int32(cap)
        _check(lib().zfv3_profile_read(self.h, names, ms, cnt, C.byref(count)))
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(count.value)}
