"""ctypes front-end of the C oracle (oracle/libgpl_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpl_oracle.so")

HD, MHD, RMHD = 1, 2, 3
PHYS = {"hd": HD, "mhd": MHD, "rmhd": RMHD}
HLL, HLLC, HLLD = 1, 2, 3
VANLEER_LIM, MC_LIM = 26, 27
SOLVERS = {"hll": HLL, "hllc": HLLC, "hlld": HLLD}
LIMITERS = {"MC_LIM": MC_LIM, "VANLEER_LIM": VANLEER_LIM, "mc": MC_LIM, "vl": VANLEER_LIM}


class Config(C.Structure):
    _fields_ = [("phys", C.c_int), ("dims", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("nz", C.c_int), ("periodic", C.c_int), ("time_stepping", C.c_int),
                ("limiter", C.c_int), ("solver", C.c_int), ("gamma", C.c_double),
                ("cfl", C.c_double), ("cfl_max_var", C.c_double),
                ("small_rho", C.c_double), ("small_prs", C.c_double), ("bc", C.c_int * 6),
                ("emf", C.c_int)]

EMF = {"UCT_HLL": 1, "UCT_HLLD": 2}

BC = {"periodic": 1, "outflow": 2, "reflective": 3}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            import subprocess
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB_PATH)
        L.gplo_create.restype = C.c_void_p
        L.gplo_create.argtypes = [C.POINTER(Config)]
        L.gplo_destroy.argtypes = [C.c_void_p]
        L.gplo_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.gplo_set_state_faces.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.gplo_get_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.gplo_advance_step.restype = C.c_uint32
        L.gplo_advance_step.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double)]
        L.gplo_run.restype = C.c_uint32
        L.gplo_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]
        L.gplo_set_cubic_math.restype = None
        L.gplo_set_cubic_math.argtypes = [C.c_void_p, C.c_void_p]
        L.gplo_zone_errors.restype = C.c_int
        L.gplo_zone_errors.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_ulonglong)]
        _lib = L
    return _lib


def set_cubic_math(acos_fn=None, sincos_fn=None):
    """Substitute the RMHD resolvent cubic's acos and sincos (C function
    addresses, e.g. ctypes.cast(lib.crm_acos, c_void_p)); no arguments
    restore the host libm, the reference's."""
    lib().gplo_set_cubic_math(acos_fn, sincos_fn)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One oracle domain.  Arrays use the periodic layout of gpluto_dev_amd.ic."""

    def __init__(self, *, phys: str, n, gamma: float, solver: str = "hlld",
                 limiter: str = "MC_LIM", time_stepping: int = 2, periodic: bool = True,
                 cfl: float = 0.4, cfl_max_var: float = 1.1, sides=None, emf: str = "UCT_HLL"):
        """sides: optional pluto.ini [Boundary] types of X1-beg, X1-end, X2-beg,
        ..., X3-end ("periodic", "outflow", "reflective"); default from
        ``periodic``."""
        nx, ny, nz = n
        dims = 3 if nz > 1 else (2 if ny > 1 else 1)
        self.n, self.dims = (nx, ny, nz), dims
        self.phys = PHYS[phys.lower()]
        self.nvar = 5 if self.phys == HD else 8
        bc = (C.c_int * 6)(*([BC[t] for t in sides] if sides else [0] * 6))
        cfg = Config(self.phys, dims, nx, ny, nz, 1 if periodic else 0, time_stepping,
                     LIMITERS[limiter], SOLVERS[solver], gamma, cfl, cfl_max_var, 1e-12, 1e-12, bc,
                     EMF[emf])
        self._h = lib().gplo_create(C.byref(cfg))
        if not self._h:
            raise ValueError("oracle: unsupported configuration")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gplo_destroy(self._h)
            self._h = None

    def set_state(self, vc, vs=None, upper=None):
        """upper: optional upper boundary face slabs per staggered direction
        (a restart dump's face index n), used where the sides are not
        periodic."""
        vc = np.ascontiguousarray(vc, dtype=np.float64)
        vs = None if vs is None else np.ascontiguousarray(vs, dtype=np.float64)
        up = None
        if upper is not None:
            up = np.ascontiguousarray(np.concatenate([np.ravel(u) for u in upper[:self.dims]]),
                                      dtype=np.float64)
        if lib().gplo_set_state_faces(self._h, _ptr(vc), _ptr(vs), _ptr(up)) != 0:
            raise ValueError("oracle: bad state")

    def get_state(self):
        nx, ny, nz = self.n
        vc = np.empty((self.nvar, nz, ny, nx))
        uc = np.empty((self.nvar, nz, ny, nx))
        vs = np.empty((self.dims, nz, ny, nx)) if self.phys != HD else None
        lib().gplo_get_state(self._h, _ptr(vc), _ptr(vs), _ptr(uc))
        return vc, vs, uc

    def advance(self, dt: float):
        inv = C.c_double(0.0)
        err = lib().gplo_advance_step(self._h, dt, C.byref(inv))
        return int(err), inv.value

    def run(self, steps: int, first_dt: float):
        t, dt = C.c_double(0.0), C.c_double(0.0)
        err = lib().gplo_run(self._h, steps, first_dt, C.byref(t), C.byref(dt))
        return int(err), t.value, dt.value

    def zone_errors(self):
        """Data::err of the active zones in each stage of the last step
        ([stage, nz, ny, nx] uint32), the zones flagged per stage, and the
        total flagged since creation."""
        nx, ny, nz = self.n
        err = np.zeros((3, nz, ny, nx), dtype=np.uint32)
        counts = np.zeros(3, dtype=np.uint64)
        total = C.c_ulonglong(0)
        ns = lib().gplo_zone_errors(self._h, _ptr(err), _ptr(counts), C.byref(total))
        return err[:ns], counts[:ns].astype(np.int64), int(total.value)
