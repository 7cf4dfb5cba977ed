"""Zone-by-zone error reports on the GPU (gpl_set_zone_reports /
gpl_get_zone_reports) against the reference's own zone-by-zone log.

The floor fixtures (tests/golden/*_floor.npz) drive ConsToPrim into its
small_prs branches (negative pressure, negative energy; Src/HD/mappers.cpp:54-90,
Src/MHD/mappers.cpp:53-96) and the Uc rewrite of flagged zones
(Src/mappers3D.cpp:182-190).  The reference's WARNING_MESSAGES = 3 build logged
every flagged zone with its step, stage and ERR_* bits (Where,
Src/debug_tools.cpp:633-682).  The GPU run, stepped as the reference's main
loop steps (Src/main.cpp:446-463), must flag exactly those zones with exactly
those bits, count them per stage as Where's err_count does
(Src/debug_tools.cpp:587-593), and end bit for bit on the reference's state."""
import numpy as np
import pytest

from golden_io import cases, expected_err, load
from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

pytestmark = pytest.mark.gpu

FLOOR = [c for c in cases() if "zones" in load(c)[1]]


def _stage(p):
    n = tuple(p["n"])
    dims = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    phys = "HD" if p["cfg"].startswith("hd") else "MHD"
    return StageUpdate(Definitions(physics=phys, dimensions=dims, time_stepping=p["time_stepping"],
                                   limiter=p["limiter"]),
                       RuntimeSetup(n=n, solver=p["solver"], cfl=p["cfl"], first_dt=p["first_dt"],
                                    gamma=p["gamma"],
                                    boundary="periodic" if p["periodic"] else "outflow",
                                    sides=tuple(p.get("sides") or ())))


def _rows(step, err):
    rows = []
    for s in range(err.shape[0]):
        k, j, i = np.nonzero(err[s])
        rows += [(step, s + 1, int(a), int(b), int(c), int(err[s, c, b, a])) for a, b, c in zip(i, j, k)]
    return rows


@pytest.mark.parametrize("name", FLOOR)
def test_zone_reports_match_reference_log(name):
    p, a = load(name)
    print(f"case {name}", flush=True)
    g = _stage(p)
    g.upload(a["vc_in"], a.get("vs_in"))
    g.set_zone_reports(True)
    dt, rows, err_all, total = p["first_dt"], [], 0, 0
    for n in range(p["steps"]):
        err, inv = g.advance_step(dt)
        err_all |= err
        zerr, counts, total = g.zone_reports()
        step_rows = _rows(n, zerr)
        rows += step_rows
        for s, c in enumerate(counts):
            assert c == sum(1 for r in step_rows if r[1] == s + 1), (n, s)
        dt = g.next_time_step(dt, inv)
    vc, vs, _ = g.download()
    g.close()
    ref = a["zones"]
    want = sorted(tuple(int(x) for x in r[:6]) for r in ref)
    assert sorted(rows) == want
    assert total == len(want)
    assert err_all == expected_err(a)
    assert np.array_equal(vc, a["vc_out"])
    if vs is not None:
        assert np.array_equal(vs, a["vs_out"])


@pytest.mark.parametrize("name", FLOOR[:1] + [c for c in FLOOR if c.startswith("mhd3d")])
def test_zone_reports_under_graph_replay(name):
    """gpl_run replays a captured step: the per-stage counters are reset by a
    kernel inside the graph, so the last step's counts and the running total
    equal the step-by-step figures; turning the reports on or off leaves the
    state untouched."""
    p, a = load(name)
    ref = a["zones"]
    g = _stage(p)
    g.upload(a["vc_in"], a.get("vs_in"))
    g.set_zone_reports(True)
    err, t, _ = g.run(p["steps"], p["first_dt"])
    zerr, counts, total = g.zone_reports()
    vc, vs, _ = g.download()
    last = p["steps"] - 1
    assert sorted(_rows(last, zerr)) == sorted(tuple(int(x) for x in r[:6]) for r in ref if r[0] == last)
    assert total == len(ref)
    for s, c in enumerate(counts):
        assert c == np.count_nonzero((ref[:, 0] == last) & (ref[:, 1] == s + 1))
    g.set_zone_reports(False)
    with pytest.raises(RuntimeError):
        g.zone_reports()
    g.close()
    assert err == expected_err(a)
    assert np.array_equal(vc, a["vc_out"])
