"""The reference-side binding (gpluto_dev_amd/bridge/gpl_bridge.cpp) compiles
against the reference's own headers for every physics module and dimension
the library drives: the definitions.hpp of oracle/ref's configurations, the
reference's Src include paths, g++ -std=c++17 (the reference's own flags).
Needs the reference sources; the linked build that runs the reference's main
loop on the GPU is tests/test_gpu_bridge.py."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/Src"
BRIDGE = os.path.join(ROOT, "gpluto_dev_amd", "bridge", "gpl_bridge.cpp")
PHYS_INC = {"hd": ["HD"], "mhd": ["MHD"], "rmhd": ["RMHD", "RHD"]}

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="needs the reference sources")


@pytest.mark.parametrize("cfg", ["hd1d", "hd2d", "hd3d", "mhd2d", "mhd3d", "rmhd3d"])
@pytest.mark.parametrize("ts,lim,emf", [("RK2", "MC_LIM", "UCT_HLL"), ("RK3", "VANLEER_LIM", "UCT_HLLD")])
def test_bridge_compiles_against_reference_headers(cfg, ts, lim, emf):
    phys = re.match(r"([a-z]+?)\d", cfg).group(1)
    inc = [f"-I{os.path.join(ROOT, 'oracle', 'ref', cfg)}", f"-I{REF}"]
    inc += [f"-I{os.path.join(REF, p)}" for p in PHYS_INC[phys]]
    inc += [f"-I{os.path.join(REF, 'States')}", f"-I{os.path.join(REF, 'Math_Tools')}"]
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror=return-type",
           f"-DGPL_TIME_STEPPING={ts}", f"-DGPL_LIMITER={lim}", f"-DGPL_CT_EMF_AVERAGE={emf}",
           *inc, BRIDGE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errors = [ln for ln in r.stderr.splitlines() if "gpl_bridge.cpp" in ln and "error" in ln]
    assert r.returncode == 0 and not errors, r.stderr[-3000:]
