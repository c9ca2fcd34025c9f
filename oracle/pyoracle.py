"""ctypes binding of the oracle (test infrastructure only).

Loads oracle/_build/liborc.so, the CPU restatement of zephyr's SmFluid step.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")

GODUNOV, RUSANOV, HLL, HLLC, HLLC_M, HLLC_LM, CIR1, CIR2 = 0, 1, 2, 3, 4, 5, 6, 7
LIMITERS = ["none", "unit", "minmod", "MC", "vanLeer", "vanAlbada1", "vanAlbada2",
            "CHARM", "HCUS", "HQUICK", "Koren", "smart", "superbee"]
UNDEFINED, INNER, WALL, ZOE, PERIODIC = -1, 0, 1, 2, 3

_lib = None

_D = C.c_double
_I = C.c_int
_P = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_flux": (_I, [_I, _dp, _dp, _dp, _dp]),
            "orc_limiter2": (_D, [_I, _D, _D]),
            "orc_limiter1": (_D, [_I, _D]),
            "orc_riemann": (None, [_D] * 10 + [_dp]),
            "orc_riemann_sample": (None, [_D] * 9 + [_dp]),
            "orc_eos": (None, [_dp, _D, _D, _D, _dp]),
            "orc_mesh_rectangle": (_P, [_I, _I, _D, _D, _D, _D, _ip]),
            "orc_mesh_rectangle_ax": (_P, [_I, _I, _D, _D, _D, _D, _ip, _I]),
            "orc_mesh_axial": (_I, [_P]),
            "orc_mesh_get_axial": (None, [_P, _dp, _dp]),
            "orc_smf_set_axial": (_I, [_P, _I]),
            "orc_smf_plant_axial": (None, [_I]),
            "orc_mesh_free": (None, [_P]),
            "orc_mesh_size": (_I, [_P]),
            "orc_mesh_get_cells": (None, [_P, _dp, _dp, _dp, _ip, _ip, _ip, _ip]),
            "orc_mesh_get_slots": (None, [_P, _ip, _ip, _dp, _dp, _dp, _dp, _dp]),
            "orc_mesh_get_verts": (None, [_P, _dp, _dp]),
            "orc_mesh_set_slots": (None, [_P, _ip, _ip, _dp, _dp, _dp, _dp, _dp]),
            "orc_mesh_incircle": (None, [_P, _dp]),
            "orc_smf_create": (_P, [_P, _D, _D, _D, _I, _I, _I, _D]),
            "orc_smf_free": (None, [_P]),
            "orc_smf_set_max_dt": (None, [_P, _D]),
            "orc_smf_set_state": (None, [_P, _dp, _dp, _dp, _dp, _dp]),
            "orc_smf_get_state": (None, [_P, _dp, _dp, _dp, _dp, _dp]),
            "orc_smf_get_wait": (None, [_P, _ip]),
            "orc_smf_set_wait": (None, [_P, _ip]),
            "orc_smf_get_grad": (None, [_P, _dp, _dp]),
            "orc_smf_update": (_I, [_P, _I]),
            "orc_smf_dt": (_D, [_P]),
            "orc_smf_compute_dt": (_D, [_P, _I]),
            "orc_smf_compute_grad": (None, [_P, _I]),
            "orc_smf_totals": (None, [_P, _dp]),
            "orc_smf_set_flags": (None, [_P, _I, _D, _I]),
            "orc_mesh_set_flags": (None, [_P, _ip]),
            "orc_mesh_balance_flags": (None, [_P, _I]),
            "orc_mesh_refine": (_I, [_P, _P, _I, _I]),
            "orc_mesh_get_next": (None, [_P, _ip]),
            "orc_splitmix_uniform": (_D, [C.POINTER(C.c_uint64)]),
            "orc_init_quadrants": (None, [_P, C.c_uint64, _D, _dp, _dp, _dp, _dp, _dp]),
            "orc_mm_create": (_P, [_P, _I, _dp, _D]),
            "orc_mm_free": (None, [_P]),
            "orc_mm_set_max_dt": (None, [_P, _D]),
            "orc_mm_set_crp": (None, [_P, _I]),
            "orc_mm_set_flux": (None, [_P, _I]),
            "orc_tr_create": (_P, [_P, _I, _D]),
            "orc_weno5": (_D, [_dp, _I]),
            "orc_tr_free": (None, [_P]),
            "orc_tr_set_velocity": (None, [_P, _dp, _dp, _dp, _dp]),
            "orc_tr_set_state": (None, [_P, _dp]),
            "orc_tr_state": (None, [_P, _dp, _dp, _dp]),
            "orc_tr_set_dt": (None, [_P, _D]),
            "orc_tr_compute_dt": (_D, [_P]),
            "orc_tr_update": (None, [_P, _I]),
            "orc_tr_slopes": (None, [_P, _dp, _dp]),
            "orc_tr_set_flags": (None, [_P]),
            "orc_mm_face_flux_kind": (_I, [_P, _I, _I, _I, _D, _D, _dp]),
            "orc_mm_stiffened_gas": (None, [_P, _I, _dp]),
            "orc_mm_sound_speed": (_D, [_P, _I]),
            "orc_mm_set_splitting": (None, [_P, _I]),
            "orc_mm_set_state": (None, [_P, _dp, _dp, _dp, _dp, _dp]),
            "orc_mm_state": (None, [_P] + [_dp] * 8),
            "orc_mm_compute_dt": (_D, [_P]),
            "orc_mm_update": (_I, [_P]),
            "orc_mm_dt": (_D, [_P]),
            "orc_mm_interface": (None, [_P, _dp, _dp, _dp]),
            "orc_mm_closure": (None, [_I, _dp, _D, _D, _dp, _D, _D, _dp]),
            "orc_mm_closure_rP": (None, [_I, _dp, _D, _D, _dp, _dp]),
            "orc_mm_face_flux": (None, [_P, _I, _I, _D, _D, _dp]),
            "orc_mm_crp_flux": (_I, [_P, _I, _I, _D, _D, _I, _D, _D, _dp]),
            "orc_mm_wave_speeds": (None, [_P, _I, _I, _dp]),
            "orc_mm_plant": (None, [_I]),
            "orc_mm_split": (None, [_P, _I, _I, _dp]),
            "orc_mm_crp_sub": (_I, [_P, _I, _I, _I, _I, _D, _D, _dp]),
            "orc_mm_set_limiter": (None, [_P, _I]),
            "orc_mm_set_accuracy": (None, [_P, _I]),
            "orc_mm_set_threads": (None, [_P, _I]),
            "orc_mm_compute_grad": (None, [_P]),
            "orc_mm_grad": (None, [_P, _dp, _dp]),
            "orc_mm_set_flags": (None, [_P]),
            "orc_mm_refine": (_I, [_P, _I]),
            "orc_quad_volume_fraction": (_D, [_D] * 5),
            "orc_quad_find_section": (_D, [_D] * 5),
            "orc_face_fraction": (_D, [_I, _D, _D]),
            "orc_average_flux": (_D, [_D] * 3),
            "orc_mesh3_cuboid": (_P, [_I, _I, _I, _D, _D, _D, _D, _D, _D, _ip]),
            "orc_mesh3_free": (None, [_P]),
            "orc_mesh3_size": (_I, [_P]),
            "orc_mesh3_get_cells": (None, [_P] + [_dp] * 5),
            "orc_mesh3_get_slots": (None, [_P, _ip, _ip] + [_dp] * 8),
            "orc_smf3_create": (_P, [_P, _D, _D, _D, _I, _I, _I, _D]),
            "orc_smf3_free": (None, [_P]),
            "orc_smf3_set_state": (None, [_P] + [_dp] * 6),
            "orc_smf3_get_state": (None, [_P] + [_dp] * 6),
            "orc_smf3_get_wait": (None, [_P, _ip]),
            "orc_smf3_set_wait": (None, [_P, _ip]),
            "orc_smf3_get_grad": (None, [_P, _dp]),
            "orc_smf3_compute_dt": (_D, [_P, _I]),
            "orc_smf3_update": (_I, [_P, _I]),
            "orc_smf3_dt": (_D, [_P]),
            "orc_smf3_totals": (None, [_P, _dp]),
            "orc_rotate3": (None, [_dp, _dp, _I, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def flux(kind, zL, zR, eos=(1.4, 0.0, 0.0)):
    """Numerical flux in the face frame; z = (rho, vx, vy, vz, p, e)."""
    zL = np.ascontiguousarray(zL, dtype=np.float64)
    zR = np.ascontiguousarray(zR, dtype=np.float64)
    e = np.ascontiguousarray(eos, dtype=np.float64)
    F = np.zeros(5)
    err = lib().orc_flux(kind, _d(zL), _d(zR), _d(e), _d(F))
    return F, err


def limiter2(kind, a, b):
    return lib().orc_limiter2(kind, a, b)


def limiter1(kind, r):
    return lib().orc_limiter1(kind, r)


def riemann(L, R, gL=1.4, p0L=0.0, gR=None, p0R=None):
    """Exact Riemann solver face state: dict(rho, Uf, Pf, P, U, conv)."""
    gR = gL if gR is None else gR
    p0R = p0L if p0R is None else p0R
    out = np.zeros(6)
    lib().orc_riemann(L[0], L[1], L[2], gL, p0L, R[0], R[1], R[2], gR, p0R, _d(out))
    return dict(rho=out[0], Uf=out[1], Pf=out[2], P=out[3], U=out[4], conv=bool(out[5]))


def riemann_sample(L, R, xi, gamma=1.4, p0=0.0):
    out = np.zeros(3)
    lib().orc_riemann_sample(L[0], L[1], L[2], R[0], R[1], R[2], gamma, p0, xi, _d(out))
    return out


def eos(params, rho, e, p):
    out = np.zeros(4)
    prm = np.ascontiguousarray(params, dtype=np.float64)
    lib().orc_eos(_d(prm), rho, e, p, _d(out))
    return dict(p=out[0], e=out[1], c=out[2], pmin=out[3])


def splitmix_stream(seed, count):
    st = C.c_uint64(seed)
    return np.array([lib().orc_splitmix_uniform(C.byref(st)) for _ in range(count)])


class Mesh:
    """Oracle AMR-2D mesh (8 face slots, 9 vertices per cell)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def rectangle(cls, nx, ny, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, bounds=(ZOE,) * 4, axial=False):
        b = np.ascontiguousarray(bounds, dtype=np.int32)
        return cls(lib().orc_mesh_rectangle_ax(nx, ny, xmin, xmax, ymin, ymax, _i(b), 1 if axial else 0))

    def axial_geometry(self):
        """(volume_alt (n), area_alt (8n)) of an axial mesh."""
        v, a = np.zeros(self.n), np.zeros(8 * self.n)
        lib().orc_mesh_get_axial(self.h, _d(v), _d(a))
        return v, a

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mesh_free(self.h)
            self.h = None

    @property
    def n(self):
        return lib().orc_mesh_size(self.h)

    def cells(self):
        n = self.n
        cx, cy, vol = np.zeros(n), np.zeros(n), np.zeros(n)
        lv, fl, bi, zi = (np.zeros(n, np.int32) for _ in range(4))
        lib().orc_mesh_get_cells(self.h, _d(cx), _d(cy), _d(vol), _i(lv), _i(fl), _i(bi), _i(zi))
        return dict(cx=cx, cy=cy, volume=vol, level=lv, flag=fl, b_idx=bi, z_idx=zi)

    def slots(self):
        ns = 8 * self.n
        b, a = np.zeros(ns, np.int32), np.zeros(ns, np.int32)
        nx, ny, fx, fy, ar = (np.zeros(ns) for _ in range(5))
        lib().orc_mesh_get_slots(self.h, _i(b), _i(a), _d(nx), _d(ny), _d(fx), _d(fy), _d(ar))
        return dict(boundary=b, adj=a, nx=nx, ny=ny, fcx=fx, fcy=fy, area=ar)

    def set_slots(self, sl):
        a = {k: np.ascontiguousarray(sl[k], dtype=np.int32 if k in ("boundary", "adj") else np.float64)
             for k in ("boundary", "adj", "nx", "ny", "fcx", "fcy", "area")}
        lib().orc_mesh_set_slots(self.h, _i(a["boundary"]), _i(a["adj"]), _d(a["nx"]), _d(a["ny"]),
                                 _d(a["fcx"]), _d(a["fcy"]), _d(a["area"]))

    def verts(self):
        nv = 9 * self.n
        vx, vy = np.zeros(nv), np.zeros(nv)
        lib().orc_mesh_get_verts(self.h, _d(vx), _d(vy))
        return vx, vy

    def incircle(self):
        h = np.zeros(self.n)
        lib().orc_mesh_incircle(self.h, _d(h))
        return h

    def set_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        lib().orc_mesh_set_flags(self.h, _i(f))

    def balance_flags(self, max_level):
        lib().orc_mesh_balance_flags(self.h, max_level)

    def refine(self, max_level, solver=None, slope=True):
        return lib().orc_mesh_refine(self.h, solver.h if solver is not None else None, max_level, int(slope))

    def quadrant_state(self, seed, gamma=1.4):
        n = self.n
        arrs = [np.zeros(n) for _ in range(5)]
        lib().orc_init_quadrants(self.h, C.c_uint64(seed), gamma, *[_d(a) for a in arrs])
        return arrs


class SmFluid:
    """Oracle SmFluid (math/solver/sm_fluid.h:22-112 restated)."""

    def __init__(self, mesh, gamma=1.4, P0=0.0, e0=0.0, flux=HLLC, limiter="MC", acc=1, cfl=0.5):
        self.mesh = mesh
        lim = LIMITERS.index(limiter) if isinstance(limiter, str) else int(limiter)
        self.h = lib().orc_smf_create(mesh.h, gamma, P0, e0, flux, lim, acc, cfl)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_smf_free(self.h)
            self.h = None

    def set_state(self, rho, u, v, p, e):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, p, e)]
        lib().orc_smf_set_state(self.h, *[_d(a) for a in arrs])

    def state(self):
        n = self.mesh.n
        arrs = [np.zeros(n) for _ in range(5)]
        lib().orc_smf_get_state(self.h, *[_d(a) for a in arrs])
        return arrs

    def wait(self):
        w = np.zeros(self.mesh.n, np.int32)
        lib().orc_smf_get_wait(self.h, _i(w))
        return w

    def set_wait(self, w):
        w = np.ascontiguousarray(w, dtype=np.int32)
        lib().orc_smf_set_wait(self.h, _i(w))

    def grad(self):
        n = self.mesh.n
        dx, dy = np.zeros(4 * n), np.zeros(4 * n)
        lib().orc_smf_get_grad(self.h, _d(dx), _d(dy))
        return dx.reshape(4, n), dy.reshape(4, n)

    def set_flags(self, criterion=0, threshold=0.1, nthreads=1):
        lib().orc_smf_set_flags(self.h, criterion, threshold, nthreads)

    def refine(self, max_level, slope=True):
        """EuMesh::refine with this solver's distributor."""
        return self.mesh.refine(max_level, self, slope)

    def set_max_dt(self, v):
        lib().orc_smf_set_max_dt(self.h, v)

    def set_axial(self, axial=True):
        """SmFluid::set_axial; the mesh must carry the axial geometry."""
        if lib().orc_smf_set_axial(self.h, 1 if axial else 0) != 0:
            raise ValueError("the mesh has no axial geometry")

    def update(self, nthreads=None):
        nt = nthreads or min(8, os.cpu_count() or 1)
        return lib().orc_smf_update(self.h, nt)

    def dt(self):
        return lib().orc_smf_dt(self.h)

    def compute_dt(self, nthreads=1):
        return lib().orc_smf_compute_dt(self.h, nthreads)

    def compute_grad(self, nthreads=1):
        lib().orc_smf_compute_grad(self.h, nthreads)

    def totals(self):
        out = np.zeros(4)
        lib().orc_smf_totals(self.h, _d(out))
        return out


# ---- three-dimensional Cuboid meshes and SmFluid -----------------------------
def rotate3(v, n, to_local=True):
    """Rotate::to_local / to_global of one 3-vector about the normal n."""
    a, b, out = np.asarray(v, np.float64), np.asarray(n, np.float64), np.zeros(3)
    lib().orc_rotate3(_d(a), _d(b), 1 if to_local else 0, _d(out))
    return out


class Mesh3:
    """Cuboid mesh of AMR-3D cells (24 face slots per cell, the six main
    ones defined): cell c = nz (ny i + j) + k; bounds L, R, B, T, Z, F."""

    def __init__(self, nx, ny, nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                 bounds=(ZOE,) * 6):
        b = np.ascontiguousarray(bounds, dtype=np.int32)
        self.shape = (nx, ny, nz)
        self.h = lib().orc_mesh3_cuboid(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, _i(b))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mesh3_free(self.h)
            self.h = None

    @property
    def n(self):
        return lib().orc_mesh3_size(self.h)

    def cells(self):
        n = self.n
        a = {k: np.zeros(n) for k in ("cx", "cy", "cz", "volume", "h")}
        lib().orc_mesh3_get_cells(self.h, *[_d(a[k]) for k in ("cx", "cy", "cz", "volume", "h")])
        return a

    def slots(self):
        ns = 24 * self.n
        a = {"boundary": np.zeros(ns, np.int32), "adj": np.zeros(ns, np.int32)}
        for k in ("nx", "ny", "nz", "fcx", "fcy", "fcz", "area", "w"):
            a[k] = np.zeros(ns)
        lib().orc_mesh3_get_slots(self.h, _i(a["boundary"]), _i(a["adj"]),
                                  *[_d(a[k]) for k in ("nx", "ny", "nz", "fcx", "fcy", "fcz", "area", "w")])
        return a


class SmFluid3:
    """Oracle SmFluid on a Cuboid mesh (oracle_3d.hpp)."""

    def __init__(self, mesh, gamma=1.4, P0=0.0, e0=0.0, flux=HLLC, limiter="MC", acc=1, cfl=0.5):
        self.mesh = mesh
        lim = LIMITERS.index(limiter) if isinstance(limiter, str) else int(limiter)
        self.h = lib().orc_smf3_create(mesh.h, gamma, P0, e0, flux, lim, acc, cfl)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_smf3_free(self.h)
            self.h = None

    def set_state(self, rho, u, v, w, p, e):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, w, p, e)]
        lib().orc_smf3_set_state(self.h, *[_d(a) for a in arrs])

    def state(self):
        arrs = [np.zeros(self.mesh.n) for _ in range(6)]
        lib().orc_smf3_get_state(self.h, *[_d(a) for a in arrs])
        return arrs

    def wait(self):
        w = np.zeros(self.mesh.n, np.int32)
        lib().orc_smf3_get_wait(self.h, _i(w))
        return w

    def set_wait(self, w):
        w = np.ascontiguousarray(w, dtype=np.int32)
        lib().orc_smf3_set_wait(self.h, _i(w))

    def grad(self):
        """d/dx, d/dy, d/dz of rho, u, v, w, p: [3][5][n]."""
        g = np.zeros(15 * self.mesh.n)
        lib().orc_smf3_get_grad(self.h, _d(g))
        return g.reshape(3, 5, self.mesh.n)

    def update(self, nthreads=None):
        nt = nthreads or min(8, os.cpu_count() or 1)
        return lib().orc_smf3_update(self.h, nt)

    def dt(self):
        return lib().orc_smf3_dt(self.h)

    def compute_dt(self, nthreads=1):
        return lib().orc_smf3_compute_dt(self.h, nthreads)

    def totals(self):
        out = np.zeros(5)
        lib().orc_smf3_totals(self.h, _d(out))
        return out


# ---- multi-material solver (MmFluid, first order, CrpMode::NONE) -----------
def _mats(materials):
    return np.ascontiguousarray(np.asarray(materials, dtype=np.float64).reshape(-1, 6))


class MmFluid:
    """Oracle MmFluid (math/solver/mm_fluid.h:45-163 restated: first or
    second order; CrpMode NONE, V3, V5, VS or PLIC; HLLC, HLL or Rusanov on
    the plain faces; direction splitting NONE, SIMPLE or STRANG); materials:
    (gamma, P0, e0, Cv, T0, rho_ref)."""

    CRP_MODES = ("NONE", "V3", "V5", "VS", "PLIC")   # CrpMode (mm_fluid.h:36-42)
    SPLITS = ("NONE", "SIMPLE", "STRANG")            # DirSplit (mm_fluid.h:29-33)

    def __init__(self, mesh, materials, cfl=0.7, crp=True):
        self.mesh = mesh
        self.mats = _mats(materials)
        self.nm = self.mats.shape[0]
        self.h = lib().orc_mm_create(mesh.h, self.nm, _d(self.mats), cfl)
        lib().orc_mm_set_crp(self.h, 4 if crp else 0)

    def set_crp_mode(self, mode):
        """CrpMode by name; composite faces always use HLLC (set_crp_mode,
        mm_fluid.cpp:63-70; NONE leaves the mode as it is there)."""
        k = self.CRP_MODES.index(mode)
        if k:
            lib().orc_mm_set_crp(self.h, k)
            lib().orc_mm_set_flux(self.h, 3)

    def set_method(self, flux):
        """set_method (mm_fluid.cpp:50-61): "CRP" = PLIC with HLLC, else
        CrpMode NONE with the named mmf flux ("HLLC", "HLL", "Rusanov",
        "Godunov")."""
        if flux == "CRP":
            lib().orc_mm_set_crp(self.h, 4)
            lib().orc_mm_set_flux(self.h, 3)
        else:
            lib().orc_mm_set_crp(self.h, 0)
            lib().orc_mm_set_flux(self.h, {"Godunov": 0, "Rusanov": 1, "HLL": 2, "HLLC": 3}[flux])

    def set_splitting(self, split):
        lib().orc_mm_set_splitting(self.h, self.SPLITS.index(split))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mm_free(self.h)
            self.h = None

    def set_max_dt(self, v):
        lib().orc_mm_set_max_dt(self.h, v)

    def set_state(self, rho, u, v, p, beta):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, p)]
        b = np.ascontiguousarray(np.asarray(beta, dtype=np.float64).reshape(self.nm, -1))
        lib().orc_mm_set_state(self.h, *[_d(a) for a in arrs], _d(b))

    def state(self):
        n, nm = self.mesh.n, self.nm
        out = {k: np.zeros(n) for k in ("rho", "u", "v", "P", "e", "T")}
        out["beta"] = np.zeros((nm, n))
        out["rhos"] = np.zeros((nm, n))
        lib().orc_mm_state(self.h, *[_d(out[k]) for k in ("rho", "u", "v", "P", "e", "T", "beta", "rhos")])
        return out

    def compute_dt(self):
        return lib().orc_mm_compute_dt(self.h)

    def update(self):
        """One MmFluid::update; returns 1 on a NaN flux or negative density."""
        return lib().orc_mm_update(self.h)

    def dt(self):
        return lib().orc_mm_dt(self.h)

    def interface(self):
        nm, n = self.nm, self.mesh.n
        nx, ny, p = np.zeros((nm, n)), np.zeros((nm, n)), np.zeros((nm, n))
        lib().orc_mm_interface(self.h, _d(nx), _d(ny), _d(p))
        return nx, ny, p

    def face_flux(self, cL, cR, nx, ny):
        F = np.zeros(9)
        lib().orc_mm_face_flux(self.h, cL, cR, nx, ny, _d(F))
        return F

    def face_flux_kind(self, kind, cL, cR, nx=1.0, ny=0.0):
        """The plain mmf flux by name ("Godunov", "Rusanov", "HLL", "HLLC")
        in the face frame, and the reference-throws flag."""
        F = np.zeros(9)
        k = {"Godunov": 0, "Rusanov": 1, "HLL": 2, "HLLC": 3}[kind]
        err = lib().orc_mm_face_flux_kind(self.h, k, cL, cR, nx, ny, _d(F))
        return F, err

    def stiffened_gas(self, c):
        """(gamma, P0) of MixturePT::stiffened_gas at cell c's state."""
        out = np.zeros(2)
        lib().orc_mm_stiffened_gas(self.h, c, _d(out))
        return out

    def sound_speed(self, c):
        return lib().orc_mm_sound_speed(self.h, c)

    def crp_flux(self, cL, cR, hL, hR, iA, a_sig, dt):
        """calc_crp_flux between cells cL and cR across a +x face: (F[9], err)."""
        F = np.zeros(9)
        err = lib().orc_mm_crp_flux(self.h, cL, cR, hL, hR, iA, a_sig, dt, _d(F))
        return F, err

    def crp_sub(self, inverse, c0, c1, c2, delta, dt):
        """CrpFlux::classic (inverse False) / inverse (True) of the states of
        cells c0, c1, c2 across a +x face: (F[9], err)."""
        F = np.zeros(9)
        err = lib().orc_mm_crp_sub(self.h, int(bool(inverse)), c0, c1, c2, delta, dt, _d(F))
        return F, err

    def split(self, c, iA):
        """PState::split of cell c for material iA: dict of the pure part's
        rho / e, the mixture part's rho / e / beta, and their common P, T."""
        out = np.zeros(10)
        lib().orc_mm_split(self.h, c, iA, _d(out))
        return {"rhoA": out[0], "eA": out[1], "rhoB": out[2], "eB": out[3], "betaB": out[4:8],
                "P": out[8], "T": out[9]}

    def wave_speeds(self, cL, cR):
        S = np.zeros(3)
        lib().orc_mm_wave_speeds(self.h, cL, cR, _d(S))
        return S

    def set_limiter(self, limiter):
        lib().orc_mm_set_limiter(self.h, LIMITERS.index(limiter) if isinstance(limiter, str) else int(limiter))

    def compute_grad(self):
        lib().orc_mm_compute_grad(self.h)

    def set_threads(self, nthreads):
        """OpenMP threads of the per-cell loops (the results do not depend on it)."""
        lib().orc_mm_set_threads(self.h, int(nthreads))

    def set_accuracy(self, acc):
        """1: the first-order update; 2: compute_grad + fluxes_stage1/2."""
        lib().orc_mm_set_accuracy(self.h, int(acc))

    def grad(self):
        """Stored gradients (gx, gy), rows rho, u, v, e, beta[n_mat], P, T,
        rhos[n_mat] (NaN where the stencil meets an absent material)."""
        K, n = 6 + 2 * self.nm, self.mesh.n
        gx, gy = np.zeros((K, n)), np.zeros((K, n))
        lib().orc_mm_grad(self.h, _d(gx), _d(gy))
        return gx, gy

    def set_flags(self):
        """MmFluid::set_flags into the mesh's flags (computes the gradients)."""
        lib().orc_mm_set_flags(self.h)

    def refine(self, max_level):
        """EuMesh::refine with MmFluid's distributor: 1 changed, 0 not, -1 a
        state the reference rejects."""
        return lib().orc_mm_refine(self.h, max_level)


def plant(k):
    """Test-only planted mistake in the oracle's composite-flux chain (0 = none)."""
    lib().orc_mm_plant(int(k))


def mm_closure(materials, rho, e, beta, P0=float("nan"), T0=float("nan")):
    """MixturePT::get_rPT: (P, T, rhos[4])."""
    m = _mats(materials)
    b = np.zeros(4)
    b[:m.shape[0]] = beta
    out = np.zeros(6)
    lib().orc_mm_closure(m.shape[0], _d(m), rho, e, _d(b), P0, T0, _d(out))
    return out[0], out[1], out[2:]


def mm_closure_rP(materials, rho, P, beta):
    """MixturePT::get_reT and sound_speed_rP: (e, T, c, rhos[4])."""
    m = _mats(materials)
    b = np.zeros(4)
    b[:m.shape[0]] = beta
    out = np.zeros(7)
    lib().orc_mm_closure_rP(m.shape[0], _d(m), rho, P, _d(b), _d(out))
    return out[0], out[1], out[2], out[3:]


def quad_volume_fraction(p, nx, ny, A, B):
    return lib().orc_quad_volume_fraction(p, nx, ny, A, B)


def quad_find_section(alpha, nx, ny, A, B):
    return lib().orc_quad_find_section(alpha, nx, ny, A, B)


FACE_FRACTION = {"plic": 0, "s": 1, "v3": 3, "v5": 5}


def face_fraction(a1, a2, kind="plic"):
    return lib().orc_face_fraction(FACE_FRACTION[kind], a1, a2)


def average_flux(alpha, cos, cfl):
    return lib().orc_average_flux(alpha, cos, cfl)


# ---- Transfer (math/solver/transfer.h) -------------------------------------
TRANSFER_METHODS = ["CRP_V3", "CRP_V5", "CRP_SE", "CRP_N1", "CRP_N2", "VOF", "VOF_CRP", "MUSCLd", "MUSCLn",
                    "MUSCLd_CRP", "MUSCLn_CRP", "MUSCL_MC", "MUSCL_MC_CRP", "WENO", "WENO_CRP"]


def weno5(u, right=True):
    """WENO5::p() (right) or m() of five cell averages u1..u5."""
    a = np.ascontiguousarray(u, dtype=np.float64)
    assert a.size == 5
    return lib().orc_weno5(_d(a), 1 if right else 0)


class Transfer:
    """Oracle Transfer: u1 carried by a velocity sampled at the face and cell
    centres (the values the reference's virtual velocity() gives there)."""

    def __init__(self, mesh, method="MUSCL_MC", cfl=0.5):
        self.mesh = mesh
        k = TRANSFER_METHODS.index(method)
        self.h = lib().orc_tr_create(mesh.h, k, cfl)
        if not self.h:
            raise ValueError(f"Transfer method {method} is not restated")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_tr_free(self.h)
            self.h = None

    def set_velocity(self, fvx, fvy, cvx, cvy):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (fvx, fvy, cvx, cvy)]
        lib().orc_tr_set_velocity(self.h, *[_d(a) for a in arrs])

    def set_state(self, u):
        lib().orc_tr_set_state(self.h, _d(np.ascontiguousarray(u, dtype=np.float64)))

    def state(self):
        n = self.mesh.n
        u, nx, ny = np.zeros(n), np.zeros(n), np.zeros(n)
        lib().orc_tr_state(self.h, _d(u), _d(nx), _d(ny))
        return u, nx, ny

    def set_dt(self, dt):
        lib().orc_tr_set_dt(self.h, dt)

    def compute_dt(self):
        return lib().orc_tr_compute_dt(self.h)

    def update(self, direction="ANY"):
        lib().orc_tr_update(self.h, ["ANY", "X", "Y"].index(direction))

    def slopes(self):
        n = self.mesh.n
        dx, dy = np.zeros(n), np.zeros(n)
        lib().orc_tr_slopes(self.h, _d(dx), _d(dy))
        return dx, dy

    def set_flags(self):
        lib().orc_tr_set_flags(self.h)
