"""The HIP path against the reference CPU build's own outputs, bit for bit
(tests/golden/*.npz written by tests/golden/make_golden.py)."""
import numpy as np
import pytest

from golden_io import cases, expected_err, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", cases("mhd3d"))
def test_gpu_matches_reference_fixture(name):
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    p, a = load(name)
    print(f"case {name}", flush=True)
    n = tuple(p["n"])
    g = StageUpdate(Definitions(time_stepping=p["time_stepping"], limiter=p["limiter"],
                                ct_emf_average=p.get("emf", "UCT_HLL")),
                    RuntimeSetup(n=n, solver=p["solver"], cfl=p["cfl"], first_dt=p["first_dt"],
                                 gamma=p["gamma"],
                                 boundary="periodic" if p["periodic"] else "outflow",
                                 sides=tuple(p.get("sides") or ())))
    g.upload(a["vc_in"], a["vs_in"])
    err, t, _ = g.run(p["steps"], p["first_dt"])
    vc, vs, _ = g.download()
    up = g.upper_faces()
    g.close()
    for d in range(3):          # outflow: the evolved boundary face; periodic: the image
        assert np.array_equal(up[d], a[f"vs_upper{d}"]), f"upper face {d}"
    assert err == expected_err(a)
    assert abs(t - p["time"]) <= 1e-6 * abs(p["time"])
    for v in range(8):
        assert np.array_equal(vc[v], a["vc_out"][v]), f"var {v}: max diff " \
            f"{np.max(np.abs(vc[v] - a['vc_out'][v])):.3e}"
    assert np.array_equal(vs, a["vs_out"])


@pytest.mark.parametrize("name", cases("mhd2d"))
def test_gpu_2d_context_matches_reference_fixture(name):
    """2D MHD through gpl_create/gpl_run (one box of the batched kernels on the
    context's stream), the upper boundary faces included."""
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    p, a = load(name)
    print(f"case {name}", flush=True)
    n = tuple(p["n"])
    g = StageUpdate(Definitions(dimensions=2, time_stepping=p["time_stepping"], limiter=p["limiter"],
                                ct_emf_average=p.get("emf", "UCT_HLL")),
                    RuntimeSetup(n=n, solver=p["solver"], cfl=p["cfl"], first_dt=p["first_dt"],
                                 gamma=p["gamma"],
                                 boundary="periodic" if p["periodic"] else "outflow",
                                 sides=tuple(p.get("sides") or ())))
    g.upload(a["vc_in"], a["vs_in"])
    err, t, _ = g.run(p["steps"], p["first_dt"])
    vc, vs, _ = g.download()
    up = g.upper_faces()
    g.close()
    assert err == expected_err(a)
    assert abs(t - p["time"]) <= 1e-6 * abs(p["time"])
    assert np.array_equal(vc, a["vc_out"]) and np.array_equal(vs, a["vs_out"])
    for d in range(2):
        assert np.array_equal(up[d], a[f"vs_upper{d}"]), f"upper face {d}"
