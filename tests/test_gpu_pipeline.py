"""The slab pipeline of the FP64 stage (sweeps of later z-slabs on one stream,
EMF and finish of completed slabs on a second, the primitives written to a
second buffer) computes exactly what the sequential kernel order computes:
bit for bit, periodic and padded boxes, MHD and RMHD, RK2 and RK3 (whose
graph replays six steps), graph replay and single steps, with the slab
count set by GPL_PIPE_SLABS at context creation."""
import os

import numpy as np
import pytest

from gpluto_dev_amd import ic
from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

pytestmark = pytest.mark.gpu


def _run(phys, n, vc, vs, steps, ts, slabs, sides=None, single=False, zones=False):
    old = os.environ.get("GPL_PIPE_SLABS")
    os.environ["GPL_PIPE_SLABS"] = str(slabs)
    try:
        g = StageUpdate(Definitions(physics=phys, time_stepping=ts),
                        RuntimeSetup(n=n, solver="hll" if phys == "RMHD" else "hlld",
                                     gamma=4 / 3 if phys == "RMHD" else 5 / 3,
                                     boundary="periodic" if not sides else "outflow",
                                     sides=tuple(sides) if sides else ()))
    finally:
        if old is None:
            del os.environ["GPL_PIPE_SLABS"]
        else:
            os.environ["GPL_PIPE_SLABS"] = old
    g.upload(vc, vs)
    if zones:
        g.set_zone_reports(True)
    if single:
        dt, err = 1e-3, 0
        for _ in range(steps):
            e, inv = g.advance_step(dt)
            err |= e
            dt = g.next_time_step(dt, inv)
        t = None
    else:
        err, t, dt = g.run(steps, 1e-3)
    v, s, u = g.download()
    z = g.zone_reports() if zones else None
    g.close()
    return err, t, dt, v, s, u, z


@pytest.mark.parametrize("phys,n,ts,slabs,sides,steps", [
    ("MHD", (64, 64, 64), "RK2", 4, None, 7),
    ("MHD", (48, 40, 96), "RK2", 8, None, 6),
    ("MHD", (40, 36, 70), "RK3", 3, None, 7),          # a short last slab; 6-step graph
    # padded box of 65 planes: slabs of 32, 32 and 1
    ("MHD", (24, 20, 60), "RK2", 3, ("outflow", "reflective", "outflow", "outflow",
                                     "reflective", "outflow"), 5),
    ("RMHD", (32, 24, 64), "RK2", 4, None, 5),
])
def test_pipeline_equals_sequential(phys, n, ts, slabs, sides, steps):
    print(f"case {phys} n={n} {ts} slabs={slabs} sides={sides}", flush=True)
    if phys == "RMHD":
        vc, vs = ic.rmhd_turbulence(n, 5, lorentz_max=3.0, sigma=1.0)
    else:
        vc, vs = ic.mhd_turbulence(n, 3)
    ref = _run(phys, n, vc, vs, steps, ts, 0, sides)
    for single in (False, True):
        got = _run(phys, n, vc, vs, steps, ts, slabs, sides, single=single)
        assert got[0] == ref[0]
        if not single:
            assert got[1] == ref[1] and got[2] == ref[2]
        for a, b in zip(got[3:6], ref[3:6]):
            assert np.array_equal(a, b)


def test_pipeline_zone_reports_equal_sequential():
    """A low-beta blast whose pressure floors flag zones in most steps."""
    n = (24, 24, 64)
    vc, vs = ic.mhd_lowbeta_blast(n, 10.0, 1e-4)
    ref = _run("MHD", n, vc, vs, 8, "RK2", 0, zones=True)
    got = _run("MHD", n, vc, vs, 8, "RK2", 4, zones=True)
    assert ref[0] != 0 and got[0] == ref[0]
    for x, y in zip(got[3:6], ref[3:6]):
        assert np.array_equal(x, y)
    for x, y in zip(got[6], ref[6]):
        assert np.array_equal(np.asarray(x), np.asarray(y))
