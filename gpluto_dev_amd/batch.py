"""Batched 2D MHD over the HIP library: many independent boxes of uneven
sizes advanced together, each with its own dt; periodic, or with outflow or
reflective sides (the same for every box of a batch).

``BatchUpdate`` mirrors ``StageUpdate`` box by box: ``upload(box, vc, vs)``,
``download(box)``, and ``run(nsteps, first_dt)`` -- the main loop of
Src/main.cpp:320-485 run for every box at once (dt of each box kept on the
device).  The reference runs one domain per process; a box here computes
what a standalone 2D run of it computes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .stage import _BOUNDARY, _EMF, _LIMITERS, _SOLVERS, _STEPPING, Definitions, RuntimeSetup


class BatchUpdate:
    def __init__(self, defs: Definitions, ini: RuntimeSetup, sizes, precision: str = "fp64",
                 device: int = 0):
        """sizes: list of N (square boxes) or of (nx, ny)."""
        L = _lib.load()
        cfg = _lib.GplConfig()
        L.gpl_default_config(C.byref(cfg))
        cfg.physics = _lib.GPL_MHD
        cfg.dimensions = 2
        for d in range(3):
            cfg.xbeg[d], cfg.xend[d] = ini.xbeg[d], ini.xend[d]
        # periodic, or outflow / reflective per side (pluto.ini [Boundary]);
        # every box of a batch has the same sides
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
        self.sizes = [(s, s) if np.isscalar(s) else tuple(s) for s in sizes]
        nb = len(self.sizes)
        nx = (C.c_int * nb)(*[s[0] for s in self.sizes])
        ny = (C.c_int * nb)(*[s[1] for s in self.sizes])
        h = C.c_void_p()
        self._L = L
        _lib.check(L.gpl_batch_create(C.byref(cfg), nb, nx, ny, C.byref(h)), "gpl_batch_create")
        self._h = h
        self.nbox = nb

    def close(self):
        if getattr(self, "_h", None):
            self._L.gpl_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, box: int, vc: np.ndarray, vs: np.ndarray,
               layout: int = _lib.GPL_LAYOUT_ACTIVE):
        """vc[8, 1, ny, nx] primitives (B3 cell-centred), vs[2, 1, ny, nx] lower faces."""
        vc = np.ascontiguousarray(vc, dtype=np.float64)
        comps = [np.ascontiguousarray(vs[d], dtype=np.float64) for d in range(2)]
        arr = (C.c_void_p * 2)(*[c.ctypes.data for c in comps])
        _lib.check(self._L.gpl_batch_upload(self._h, box, vc.ctypes.data, arr, layout),
                   "gpl_batch_upload")

    def download(self, box: int):
        nx, ny = self.sizes[box]
        vc = np.empty((8, 1, ny, nx))
        uc = np.empty((8, 1, ny, nx))
        vs = np.empty((2, 1, ny, nx))
        arr = (C.c_void_p * 2)(*[vs[d].ctypes.data for d in range(2)])
        _lib.check(self._L.gpl_batch_download(self._h, box, vc.ctypes.data, arr, uc.ctypes.data,
                                              _lib.GPL_LAYOUT_ACTIVE), "gpl_batch_download")
        return vc, vs, uc

    def upper_faces(self, box: int):
        """The box's upper boundary faces (index n along x and y), slabs
        [1, ny, 1] and [1, 1, nx] as the reference's Vs stores them: the
        periodic image of the lower face on a periodic side, the evolved face
        on an outflow or reflective side."""
        nx, ny = self.sizes[box]
        tx, ty = nx + 4, ny + 4          # Ary4D with gpl_config.nghost = 2 along x and y
        arrs = [np.zeros((1, ty, tx + 1)), np.zeros((1, ty + 1, tx))]
        ptrs = (C.c_void_p * 2)(*[a.ctypes.data for a in arrs])
        _lib.check(self._L.gpl_batch_download(self._h, box, None, ptrs, None, _lib.GPL_LAYOUT_ARY4D),
                   "gpl_batch_download")
        return [arrs[0][:, 2:2 + ny, 2 + nx:3 + nx], arrs[1][:, 2 + ny:3 + ny, 2:2 + nx]]

    def run(self, nsteps: int, first_dt=None):
        """Returns (err, time[nbox], dt[nbox]); first_dt: scalar, per-box array or None."""
        t = np.zeros(self.nbox)
        dt = np.zeros(self.nbox)
        fd = None
        if first_dt is not None:
            fd = np.ascontiguousarray(np.broadcast_to(np.asarray(first_dt, dtype=np.float64),
                                                      (self.nbox,)))
        dp = C.POINTER(C.c_double)
        err = self._L.gpl_batch_run(self._h, nsteps,
                                    None if fd is None else fd.ctypes.data_as(dp),
                                    t.ctypes.data_as(dp), dt.ctypes.data_as(dp))
        return _lib.check_step(err, "gpl_batch_run"), t, dt

    def set_tuning(self, march: int, blocks_per_cu: int = 0):
        _lib.check(self._L.gpl_batch_set_tuning(self._h, march, blocks_per_cu),
                   "gpl_batch_set_tuning")

    def tuning(self):
        m, p, ms = C.c_int(0), C.c_int(0), C.c_double(0.0)
        self._L.gpl_batch_get_tuning(self._h, C.byref(m), C.byref(p), C.byref(ms))
        return m.value, p.value, ms.value

    def last_run_ms(self) -> float:
        return self._L.gpl_batch_last_run_ms(self._h)

    def device_bytes(self) -> int:
        return int(self._L.gpl_batch_device_bytes(self._h))
