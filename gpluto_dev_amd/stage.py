"""Host-side mirror of the reference's time-step seam over the HIP library.

``StageUpdate`` holds one device domain and exposes the reference's entry
points with the same meaning:

* ``advance_step(dt)``        -- AdvanceStep(data, grid, Dts)   (Src/rk_step.cpp:88)
* ``next_time_step(dt, inv)`` -- NextTimeStep(Dts, ...)          (Src/next_time_step.cpp:20)
* ``run(nsteps, first_dt)``   -- the main loop of Src/main.cpp:320-485 (dt kept on the device)

Selections come from the same two places as in the reference: the
definitions.hpp macros (``Definitions``, or ``parse_definitions``) and the
pluto.ini sections (``RuntimeSetup``, or ``parse_pluto_ini``).  Errors are
returned as the reference's ERR_* bitmask (Src/gpluto.hpp:247-265).
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import _lib

ERR_NEGATIVE_RHO = 1 << 0
ERR_NEGATIVE_ENG = 1 << 1
ERR_NEGATIVE_PRS = 1 << 2
ERR_CONS2PRIM_FAIL = 1 << 4
ERR_NAN = 1 << 11
ERR_QUIT = 1 << 15

_LIMITERS = {"MC_LIM": _lib.GPL_MC_LIM, "VANLEER_LIM": _lib.GPL_VANLEER_LIM,
             "DEFAULT": _lib.GPL_VANLEER_LIM}
_SOLVERS = {"hlld": _lib.GPL_SOLVER_HLLD, "hll": _lib.GPL_SOLVER_HLL, "hllc": _lib.GPL_SOLVER_HLLC}
_STEPPING = {"RK2": _lib.GPL_RK2, "RK3": _lib.GPL_RK3}
_EMF = {"UCT_HLL": _lib.GPL_UCT_HLL, "UCT_HLLD": _lib.GPL_UCT_HLLD}
_BOUNDARY = {"periodic": _lib.GPL_PERIODIC, "outflow": _lib.GPL_OUTFLOW,
             "reflective": _lib.GPL_REFLECTIVE}


@dataclass
class Definitions:
    """definitions.hpp selections the stage update reads (Src/gpluto.hpp:385-561)."""
    physics: str = "MHD"
    dimensions: int = 3
    time_stepping: str = "RK2"
    limiter: str = "MC_LIM"                  # default VANLEER_LIM (Src/gpluto.hpp:462-464)
    ct_emf_average: str = "UCT_HLL"          # default (Src/MHD/CT/ct.hpp:89-96)
    reconstruction: str = "LINEAR"


@dataclass
class RuntimeSetup:
    """pluto.ini selections (Src/runtime_setup.cpp:34-160)."""
    n: tuple = (64, 64, 64)
    xbeg: tuple = (0.0, 0.0, 0.0)
    xend: tuple = (1.0, 1.0, 1.0)
    cfl: float = 0.4
    cfl_max_var: float = 1.1
    first_dt: float = 1.e-4
    solver: str = "hlld"
    boundary: str = "periodic"
    sides: tuple = ()        # X1-beg, X1-end, X2-beg, X2-end, X3-beg, X3-end when they differ
    gamma: float = 5.0 / 3.0
    extra: dict = field(default_factory=dict)


def parse_definitions(text: str) -> Definitions:
    d = Definitions()
    macros = dict(re.findall(r"^\s*#define\s+(\w+)\s+(\S+)", text, flags=re.M))
    d.physics = macros.get("PHYSICS", d.physics)
    d.dimensions = int(macros.get("DIMENSIONS", d.dimensions))
    d.time_stepping = macros.get("TIME_STEPPING", d.time_stepping)
    d.limiter = macros.get("LIMITER", "VANLEER_LIM")
    d.ct_emf_average = macros.get("CT_EMF_AVERAGE", d.ct_emf_average)
    d.reconstruction = macros.get("RECONSTRUCTION", d.reconstruction)
    return d


def parse_pluto_ini(text: str) -> RuntimeSetup:
    r = RuntimeSetup()
    kv = {}
    for line in text.splitlines():
        line = line.split("#")[0].strip()
        if not line or line.startswith("["):
            continue
        parts = line.split()
        kv[parts[0]] = parts[1:]
    n, xb, xe = [], [], []
    for ax in ("X1-grid", "X2-grid", "X3-grid"):
        f = kv.get(ax, ["0.0", "1", "1.0"])
        # legacy 3-field form "xbeg n xend" (Src/grid_input.hpp)
        xb.append(float(f[0]))
        n.append(int(f[1]))
        xe.append(float(f[2]))
    r.n, r.xbeg, r.xend = tuple(n), tuple(xb), tuple(xe)
    r.cfl = float(kv.get("CFL", [r.cfl])[0])
    r.cfl_max_var = float(kv.get("CFL_max_var", [r.cfl_max_var])[0])
    r.first_dt = float(kv.get("first_dt", [r.first_dt])[0])
    r.solver = kv.get("Solver", [r.solver])[0]
    r.boundary = kv.get("X1-beg", [r.boundary])[0]
    r.sides = tuple(kv.get(f"{ax}-{e}", [r.boundary])[0]
                    for ax in ("X1", "X2", "X3") for e in ("beg", "end"))
    return r


class StageUpdate:
    """One device domain of the Godunov stage update."""

    def __init__(self, defs: Definitions, ini: RuntimeSetup, precision: str = "fp64",
                 device: int = 0):
        L = _lib.load()
        cfg = _lib.GplConfig()
        L.gpl_default_config(C.byref(cfg))
        cfg.physics = {"MHD": _lib.GPL_MHD, "RMHD": _lib.GPL_RMHD, "HD": _lib.GPL_HD}[defs.physics.upper()]
        cfg.dimensions = defs.dimensions
        cfg.nx, cfg.ny, cfg.nz = ini.n
        for d in range(3):
            cfg.xbeg[d], cfg.xend[d] = ini.xbeg[d], ini.xend[d]
        cfg.boundary = _BOUNDARY[ini.boundary]
        for q, t in enumerate(ini.sides):
            cfg.side_bc[q] = _BOUNDARY[t]
        cfg.time_stepping = _STEPPING[defs.time_stepping]
        cfg.limiter = _LIMITERS[defs.limiter]
        cfg.solver = _SOLVERS[ini.solver]
        cfg.ct_emf_average = _EMF[defs.ct_emf_average]
        cfg.precision = _lib.GPL_FP64 if precision == "fp64" else _lib.GPL_MIXED
        cfg.device = device
        cfg.gamma = ini.gamma
        cfg.cfl = ini.cfl
        cfg.cfl_max_var = ini.cfl_max_var
        self.defs, self.ini, self.precision = defs, ini, precision
        self._L = L
        h = C.c_void_p()
        _lib.check(L.gpl_create(C.byref(cfg), C.byref(h)), "gpl_create")
        self._h = h
        self.n = tuple(ini.n)
        self.mhd = cfg.physics != _lib.GPL_HD   # staggered field + CT (MHD, RMHD)
        self.nvar = 8 if self.mhd else 5      # NVAR (Src/MHD/mhd.hpp:21-41, Src/HD/hd.hpp:21-34)
        self._nfaces = defs.dimensions if self.mhd else 0   # staggered components

    def close(self):
        if getattr(self, "_h", None):
            self._L.gpl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state transfer -----------------------------------------------------
    def upload(self, vc: np.ndarray, vs: np.ndarray | None = None,
               layout: int = _lib.GPL_LAYOUT_ACTIVE):
        """vc[NVAR, nz, ny, nx] primitives; MHD: vs[DIMENSIONS, nz, ny, nx] lower
        faces (2D: B1, B2 on faces, B3 cell-centred in vc)."""
        vc = np.ascontiguousarray(vc, dtype=np.float64)
        if vc.shape[0] != self.nvar:
            raise ValueError(f"vc holds {vc.shape[0]} variables, expected {self.nvar}")
        arr = None
        if self.mhd:
            comps = [np.ascontiguousarray(vs[d], dtype=np.float64) for d in range(self._nfaces)]
            arr = (C.c_void_p * 3)(*[c.ctypes.data for c in comps], *[None] * (3 - self._nfaces))
        _lib.check(self._L.gpl_upload(self._h, vc.ctypes.data, arr, layout), "gpl_upload")

    def download(self):
        """(vc, vs, uc) on the active zones; vs is None for HD."""
        nx, ny, nz = self.n
        vc = np.empty((self.nvar, nz, ny, nx))
        uc = np.empty((self.nvar, nz, ny, nx))
        vs, arr = None, None
        if self.mhd:
            vs = np.empty((self._nfaces, nz, ny, nx))
            arr = (C.c_void_p * 3)(*[vs[d].ctypes.data for d in range(self._nfaces)],
                                   *[None] * (3 - self._nfaces))
        _lib.check(self._L.gpl_download(self._h, vc.ctypes.data, arr, uc.ctypes.data,
                                        _lib.GPL_LAYOUT_ACTIVE), "gpl_download")
        return vc, vs, uc

    def upper_faces(self):
        """The staggered field's upper boundary faces (index n along each
        direction), slabs [nz, ny, 1], [nz, 1, nx], [1, ny, nx] as the
        reference's Vs stores them: the periodic image of the lower face on a
        periodic side, the evolved face on an outflow side."""
        nx, ny, nz = self.n
        # the Ary4D ghost zones (gpl_config.nghost = 2) along the used directions
        gx, gy, gz = (2 if self.defs.dimensions > d else 0 for d in range(3))
        t = (nx + 2 * gx, ny + 2 * gy, nz + 2 * gz)
        arrs = [np.zeros((t[2] + (d == 2), t[1] + (d == 1), t[0] + (d == 0))) for d in range(3)]
        ptrs = (C.c_void_p * 3)(*[a.ctypes.data for a in arrs])
        _lib.check(self._L.gpl_download(self._h, None, ptrs, None, _lib.GPL_LAYOUT_ARY4D),
                   "gpl_download")
        slabs = [arrs[0][gz:gz + nz, gy:gy + ny, gx + nx:gx + nx + 1],
                 arrs[1][gz:gz + nz, gy + ny:gy + ny + 1, gx:gx + nx],
                 arrs[2][gz + nz:gz + nz + 1, gy:gy + ny, gx:gx + nx]]
        return slabs[:self._nfaces]          # 2D MHD: B1 and B2

    # -- step state, dumps and restarts --------------------------------------
    def set_step(self, time: float, dt: float, nstep: int):
        """timeStep time / dt and g_stepNumber of the device loop."""
        _lib.check(self._L.gpl_set_step(self._h, time, dt, nstep), "gpl_set_step")

    def get_step(self):
        """(time, dt, nstep) of the device loop."""
        t, dt, n = C.c_double(0.0), C.c_double(0.0), C.c_longlong(0)
        _lib.check(self._L.gpl_get_step(self._h, C.byref(t), C.byref(dt), C.byref(n)), "gpl_get_step")
        return t.value, dt.value, n.value

    def write_dbl(self, directory: str, nfile: int, nstep: int | None = None):
        """WriteData (dbl, single_file) + RestartDump: data.<nfile>.dbl, line
        nfile of dbl.out, record nfile of restart.out in `directory`; the
        step counter written is `nstep`, or the device's own."""
        _lib.check(self._L.gpl_write_dbl(self._h, os.fsencode(directory), nfile,
                                         -1 if nstep is None else nstep), "gpl_write_dbl")

    def write_dbl_async(self, directory: str, nfile: int, nstep: int | None = None):
        """write_dbl overlapped with later calls: the state is snapshot on the
        device now and written by a host thread; wait_dbl() joins it."""
        _lib.check(self._L.gpl_write_dbl_async(self._h, os.fsencode(directory), nfile,
                                               -1 if nstep is None else nstep), "gpl_write_dbl_async")

    def wait_dbl(self):
        _lib.check(self._L.gpl_wait_dbl(self._h), "gpl_wait_dbl")

    def read_dbl(self, directory: str, nfile: int):
        """RestartFromFile: the state, time, dt and step counter of dump nfile."""
        _lib.check(self._L.gpl_read_dbl(self._h, os.fsencode(directory), nfile), "gpl_read_dbl")

    # -- stepping -----------------------------------------------------------
    def advance_step(self, dt: float):
        inv = C.c_double(0.0)
        err = _lib.check_step(self._L.gpl_advance_step(self._h, dt, C.byref(inv)), "gpl_advance_step")
        return err, inv.value

    def next_time_step(self, dt: float, inv_dt: float) -> float:
        return self._L.gpl_next_time_step(self._h, dt, inv_dt)

    def run(self, nsteps: int, first_dt: float):
        t, dt = C.c_double(0.0), C.c_double(0.0)
        err = _lib.check_step(self._L.gpl_run(self._h, nsteps, first_dt, C.byref(t), C.byref(dt)),
                              "gpl_run")
        return err, t.value, dt.value

    def set_c2p_network(self, weights=None):
        """RMHD: learned con2prim guess (gpluto_dev_amd.c2p_net.train); None
        restores the reference's Eq. (A27) start."""
        if weights is None:
            _lib.check(self._L.gpl_set_c2p_network(self._h, None), "gpl_set_c2p_network")
            return
        w = np.ascontiguousarray(weights, dtype=np.float32)
        _lib.check(self._L.gpl_set_c2p_network(self._h, w.ctypes.data_as(C.POINTER(C.c_float))),
                   "gpl_set_c2p_network")

    def diagnostics(self) -> dict:
        """DataMinMax / CT_CheckDivB of the current state, reduced on the device:
        {"min": [NVAR], "max": [NVAR], "mach_max": float, "divb_max": float}."""
        vmin = np.empty(self.nvar)
        vmax = np.empty(self.nvar)
        mach, divb = C.c_double(0.0), C.c_double(0.0)
        dp = C.POINTER(C.c_double)
        _lib.check(self._L.gpl_diagnostics(self._h, vmin.ctypes.data_as(dp), vmax.ctypes.data_as(dp),
                                           C.byref(mach), C.byref(divb)), "gpl_diagnostics")
        return {"min": vmin, "max": vmax, "mach_max": mach.value, "divb_max": divb.value}

    def set_zone_reports(self, on: bool = True):
        """Zone-by-zone error reports (the reference's Data::err and the
        flagged-zone counts of Where): off by default."""
        _lib.check(self._L.gpl_set_zone_reports(self._h, 1 if on else 0), "gpl_set_zone_reports")

    def zone_reports(self):
        """(err[stage, nz, ny, nx] uint32 of the last step, flagged zones per
        stage of the last step, flagged zones since the reports were enabled)."""
        nx, ny, nz = self.n
        err = np.zeros((3, nz, ny, nx), dtype=np.uint32)
        counts = np.zeros(3, dtype=np.uint64)
        total = C.c_ulonglong(0)
        ns = self._L.gpl_get_zone_reports(self._h, err.ctypes.data_as(C.c_void_p),
                                          counts.ctypes.data_as(C.c_void_p), C.byref(total))
        _lib.check(min(ns, 0), "gpl_get_zone_reports")
        return err[:ns], counts[:ns].astype(np.int64), int(total.value)

    # -- measurement ----------------------------------------------------------
    def set_timing(self, on: bool):
        _lib.check(self._L.gpl_set_timing(self._h, 1 if on else 0), "gpl_set_timing")

    def kernel_stats(self) -> dict:
        out = {}
        for k in range(self._L.gpl_kernel_count(self._h)):
            name, cnt, ms = C.c_char_p(), C.c_longlong(0), C.c_double(0.0)
            self._L.gpl_kernel_stats(self._h, k, C.byref(name), C.byref(cnt), C.byref(ms))
            out[name.value.decode()] = (cnt.value, ms.value)
        return out

    def last_run_ms(self) -> float:
        return self._L.gpl_last_run_ms(self._h)

    def device_bytes(self) -> int:
        return int(self._L.gpl_device_bytes(self._h))
