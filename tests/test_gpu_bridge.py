"""The reference's own main loop with AdvanceStep on the GPU.

oracle/ref/Makefile (BRIDGE=1) links the reference's objects without
rk_step.o, plus gpluto_dev_amd/bridge/gpl_bridge.cpp and libgpluto_amd.so:
the reference's main (Src/main.cpp) reads pluto.ini, sets up the IC through
its own Init/Startup, calls AdvanceStep at Src/main.cpp:446 -- now the
library -- and NextTimeStep, and writes its data.NNNN.dbl from its host
arrays.  That file must equal the fixture the unmodified reference build
wrote, bit for bit (periodic, outflow and wall sides; MHD in 3D and 2D, HD;
RMHD within the math-library bound of tests/test_gpu_rmhd.py)."""
import os

import numpy as np
import pytest

from golden_io import load

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _binary(cfg):
    name = {"mhd3d": "mhd3d_RK2_MC_LIM_UCT_HLL", "mhd2d": "mhd2d_RK2_MC_LIM_UCT_HLL",
            "hd1d": "hd1d_RK2_MC_LIM", "rmhd3d": "rmhd3d_RK2_MC_LIM_UCT_HLL"}[cfg]
    path = os.path.join(ROOT, "oracle", "_ref", name + "_bridge", "pluto")
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: build it with make -C oracle/ref CFG={cfg} BRIDGE=1")
    return path


@pytest.mark.parametrize("name", ["mhd3d_16_rk2_mc", "mhd3d_12x14x10_rk2_mc_outflow",
                                  "mhd3d_12x10x14_rk2_mc_walls", "mhd3d_12x14x10_rk2_mc_floor",
                                  "mhd2d_32x24_rk2_mc", "hd1d_4096_rk2_mc_hllc",
                                  "hd1d_256_rk2_mc_hllc_floor", "rmhd3d_16_rk2_mc_hll"])
def test_reference_main_loop_drives_the_gpu(name):
    import ref_run
    p, a = load(name)
    print(f"case {name}", flush=True)
    r = ref_run.run_reference(_binary(p["cfg"]), a["vc_in"], a.get("vs_in"), gamma=p["gamma"],
                              steps=p["steps"], cfl=p["cfl"], first_dt=p["first_dt"],
                              solver=p["solver"], periodic=p["periodic"], sides=p.get("sides"),
                              timeout=300)
    assert abs(r.time - p["time"]) <= 1e-6 * abs(p["time"])
    if p["cfg"].startswith("rmhd"):
        for v in range(8):
            den = np.sum(np.abs(a["vc_out"][v]))
            assert np.sum(np.abs(r.vc[v] - a["vc_out"][v])) <= 1e-12 * den, v
        return
    assert np.array_equal(r.vc, a["vc_out"])
    if "vs_out" in a:
        assert np.array_equal(r.vs, a["vs_out"])
        for d in range(len(r.vs_upper)):
            assert np.array_equal(r.vs_upper[d], a[f"vs_upper{d}"]), f"upper face {d}"
