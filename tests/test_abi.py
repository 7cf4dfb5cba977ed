"""The C-ABI library loads and exports every function include/gpluto_amd.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(gpl_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    for must in ("gpl_create", "gpl_destroy", "gpl_upload", "gpl_download",
                 "gpl_advance_step", "gpl_next_time_step", "gpl_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from gpluto_dev_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libgpluto_amd.so not built (python -m gpluto_dev_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(declared())


def test_default_config_and_strerror_without_gpu():
    from gpluto_dev_amd import _lib
    L = _lib.load()
    cfg = _lib.GplConfig()
    L.gpl_default_config(ctypes.byref(cfg))
    assert (cfg.nx, cfg.limiter, cfg.solver, cfg.time_stepping) == (64, 27, 3, 2)
    assert abs(cfg.gamma - 5 / 3) < 1e-15 and cfg.cfl == 0.4 and cfg.cfl_max_var == 1.1
    assert L.gpl_strerror(_lib.GPL_EINVAL) == b"invalid argument"


def test_config_parsers_follow_reference_files():
    from gpluto_dev_amd.stage import parse_definitions, parse_pluto_ini
    d = parse_definitions(open(os.path.join(ROOT, "oracle", "ref", "mhd3d", "definitions.hpp")).read()
                          .replace("GPL_TIME_STEPPING", "RK2").replace("GPL_LIMITER", "MC_LIM")
                          .replace("GPL_CT_EMF_AVERAGE", "UCT_HLL"))
    assert (d.physics, d.dimensions, d.time_stepping, d.limiter) == ("MHD", 3, "RK2", "MC_LIM")
    ini = parse_pluto_ini("[Grid]\nX1-grid 0.0 256 1.0\nX2-grid 0.0 128 1.0\nX3-grid 0.0 64 1.0\n"
                          "[Time]\nCFL 0.4\nCFL_max_var 1.1\nfirst_dt 1.e-4\n[Solver]\nSolver hlld\n"
                          "[Boundary]\nX1-beg periodic\n")
    assert ini.n == (256, 128, 64) and ini.solver == "hlld" and ini.cfl == 0.4
