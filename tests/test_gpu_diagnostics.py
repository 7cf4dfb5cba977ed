"""On-device diagnostics (DataMinMax, Src/main.cpp:752-860; CT_CheckDivB,
Src/MHD/CT/ct_update.cpp:580-635) against the downloaded state.

Min and max are exact whatever the reduction order, so they must equal the
host's to the bit; the Mach number is a per-cell IEEE expression, so its
maximum must too.  div B is pinned independently of its formula: the seeded
IC is divergence-free to round-off, and raising one x-face by delta makes the
two cells on either side of it diverge by exactly the face area times delta.
"""
import numpy as np
import pytest

from gpluto_dev_amd import ic
from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

pytestmark = pytest.mark.gpu


def _mach(vc, prs):
    cs = np.sqrt(vc[prs] / vc[0])
    return float(np.max(np.maximum(np.maximum(np.abs(vc[1]) / cs, np.abs(vc[2]) / cs),
                                   np.abs(vc[3]) / cs)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_min_max_mach_match_the_downloaded_state(precision):
    n = (24, 20, 16)
    print(f"case diagnostics mhd n={n} {precision}", flush=True)
    vc, vs = ic.mhd_turbulence(n, 5)
    g = StageUpdate(Definitions(), RuntimeSetup(n=n), precision=precision)
    g.upload(vc, vs)
    err, _, _ = g.run(4, 1e-3)
    assert err == 0
    d = g.diagnostics()
    v, s, _ = g.download()
    g.close()
    for q in range(8):
        assert d["min"][q] == v[q].min() and d["max"][q] == v[q].max(), q
    assert d["mach_max"] == _mach(v, 7)
    assert d["divb_max"] <= 1e-12 * np.max(np.abs(s))


def test_divb_of_a_planted_face_perturbation():
    n = (16, 12, 10)
    print(f"case diagnostics divb n={n}", flush=True)
    vc, vs = ic.mhd_turbulence(n, 6)
    g = StageUpdate(Definitions(), RuntimeSetup(n=n))
    g.upload(vc, vs)
    base = g.diagnostics()["divb_max"]
    assert base <= 1e-12 * np.max(np.abs(vs))
    delta = 0.25
    vs2 = vs.copy()
    i, j, k = 5, 7, 3
    vs2[0, k, j, i] += delta               # lower x-face of cell (i, j, k)
    g.upload(vc, vs2)
    d = g.diagnostics()
    g.close()
    nx, ny, nz = n
    ax = 1.0 * (1.0 / ny) * (1.0 / nz)     # dy*dz on the unit cube
    # cells (i-1, j, k) and (i, j, k) see +-ax*delta on top of round-off
    assert abs(d["divb_max"] - ax * delta) <= 1e-10 * ax * delta + 2 * base


def test_2d_mhd_diagnostics():
    """2D MHD (one box of the batched kernels behind gpl_create): min/max/Mach
    of the downloaded state, div B = dBx1 + dBx2 with the 2D face areas
    (no z term), on a box with walls and outflow sides."""
    n = (30, 22, 1)
    print(f"case diagnostics mhd2d n={n}", flush=True)
    vc, vs = ic.mhd_turbulence(n, 8)
    sides = ("reflective", "outflow", "outflow", "reflective", "outflow", "outflow")
    g = StageUpdate(Definitions(dimensions=2), RuntimeSetup(n=n, sides=sides))
    g.upload(vc, vs)
    err, _, _ = g.run(4, 1e-3)
    assert err == 0
    d = g.diagnostics()
    v, s, _ = g.download()
    for q in range(8):
        assert d["min"][q] == v[q].min() and d["max"][q] == v[q].max(), q
    assert d["mach_max"] == _mach(v, 7)
    assert d["divb_max"] <= 1e-12 * np.max(np.abs(s))
    # the initial field with one face raised (an active-layout upload starts
    # the upper boundary faces as images of face 0, which the periodic
    # initial field satisfies)
    g.upload(vc, vs)
    base = g.diagnostics()["divb_max"]
    assert base <= 1e-12 * np.max(np.abs(vs))
    delta = 0.25
    vs2 = vs.copy()
    i, j = 9, 13
    vs2[1, 0, j, i] += delta               # lower y-face of cell (i, j)
    g.upload(vc, vs2)
    d2 = g.diagnostics()
    g.close()
    ay = (1.0 / n[0]) * 1.0                # dx*1.0 on the unit square
    # cells (i, j-1) and (i, j) see -+ay*delta on top of round-off
    assert abs(d2["divb_max"] - ay * delta) <= 1e-10 * ay * delta + 2 * base


def test_hd_diagnostics():
    n = 4096
    print(f"case diagnostics hd n={n}", flush=True)
    vc = ic.hd_riemann(n, 3)
    g = StageUpdate(Definitions(physics="HD", dimensions=1),
                    RuntimeSetup(n=(n, 1, 1), solver="hllc", cfl=0.8, gamma=1.4, boundary="outflow"))
    g.upload(vc)
    err, _, _ = g.run(50, 1e-4)
    assert err == 0
    d = g.diagnostics()
    v, _, _ = g.download()
    g.close()
    for q in range(5):
        assert d["min"][q] == v[q].min() and d["max"][q] == v[q].max(), q
    assert d["mach_max"] == _mach(v, 4)
    assert d["divb_max"] == 0.0
