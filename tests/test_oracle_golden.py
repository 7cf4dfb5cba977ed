"""The oracle against the reference CPU build's own outputs, bit for bit.

Fixtures: tests/golden/*.npz, written by tests/golden/make_golden.py from the
reference binary built from its sources (oracle/ref/Makefile).
"""
import re

import numpy as np
import pytest

from golden_io import cases, load


@pytest.mark.parametrize("name", cases())
def test_oracle_matches_reference_bitwise(name):
    import oracle
    p, a = load(name)
    phys = re.match(r"([a-z]+?)\d", p["cfg"]).group(1)    # mhd3d -> mhd, rmhd3d -> rmhd
    o = oracle.Oracle(phys=phys, n=tuple(p["n"]), gamma=p["gamma"], solver=p["solver"],
                      limiter=p["limiter"], time_stepping=int(p["time_stepping"][-1]),
                      periodic=p["periodic"], cfl=p["cfl"], cfl_max_var=p["cfl_max_var"],
                      sides=p.get("sides"), emf=p.get("emf", "UCT_HLL"))
    o.set_state(a["vc_in"], a.get("vs_in"))
    err, t, _ = o.run(p["steps"], p["first_dt"])
    vc, vs, _ = o.get_state()
    # a case may drive cells through the reference's recovery branches (RMHD
    # pressure fix); never into ERR_QUIT
    assert err & (1 << 15) == 0 and (err == 0) == (not p.get("err_log", False))
    # dbl.out prints the time with 7 significant digits
    assert abs(t - p["time"]) <= 1e-6 * abs(p["time"])
    for v in range(vc.shape[0]):
        assert np.array_equal(vc[v], a["vc_out"][v]), f"{name}: var {v} max diff " \
            f"{np.max(np.abs(vc[v] - a['vc_out'][v])):.3e}"
    if "vs_out" in a:
        assert np.array_equal(vs, a["vs_out"])


@pytest.mark.parametrize("name", [c for c in cases("mhd") + cases("rmhd") if load(c)[0]["periodic"]])
def test_reference_duplicate_faces_are_periodic_images(name):
    """The reference's extra upper face equals the lower face bitwise, which is
    what lets a ghost-free periodic layout reproduce it exactly."""
    p, a = load(name)
    vs = a["vs_out"]
    for d in range(vs.shape[0]):
        up = a[f"vs_upper{d}"]
        lo = [vs[d][:, :, :1], vs[d][:, :1, :], vs[d][:1]][d]
        assert np.array_equal(up, lo)


def zone_rows(step, err):
    """(step, stage, i, j, k, bits) rows of the flagged zones in err[stage, k, j, i]."""
    rows = []
    for s in range(err.shape[0]):
        k, j, i = np.nonzero(err[s])
        rows += [(step, s + 1, int(a), int(b), int(c), int(err[s, c, b, a])) for a, b, c in zip(i, j, k)]
    return rows


def run_with_zone_reports(o, steps, first_dt, cfl, cfl_max_var):
    """Step the oracle as the reference's main loop does (Src/main.cpp:446-463)
    and collect every stage's flagged zones and the per-stage counts."""
    dt, rows, counts = first_dt, [], []
    for n in range(steps):
        _, inv = o.advance(dt)
        err, cnt, _ = o.zone_errors()
        rows += zone_rows(n, err)
        counts.append(cnt)
        dt = min(cfl / inv, cfl_max_var * dt)
    return rows, counts


@pytest.mark.parametrize("name", [c for c in cases() if "zones" in load(c)[1]])
def test_oracle_zone_reports_match_reference(name):
    """Every zone the reference's ConsToPrim3D flags (its WARNING_MESSAGES = 3
    log: step, stage, zone and ERR_* bits) is flagged by the oracle with the
    same bits, and no other; the per-stage counts are Where's err_count
    increments (Src/debug_tools.cpp:587-593)."""
    import oracle
    p, a = load(name)
    phys = re.match(r"([a-z]+?)\d", p["cfg"]).group(1)
    o = oracle.Oracle(phys=phys, n=tuple(p["n"]), gamma=p["gamma"], solver=p["solver"],
                      limiter=p["limiter"], time_stepping=int(p["time_stepping"][-1]),
                      periodic=p["periodic"], cfl=p["cfl"], cfl_max_var=p["cfl_max_var"],
                      sides=p.get("sides"), emf=p.get("emf", "UCT_HLL"))
    o.set_state(a["vc_in"], a.get("vs_in"))
    rows, counts = run_with_zone_reports(o, p["steps"], p["first_dt"], p["cfl"], p["cfl_max_var"])
    ref = a["zones"]
    assert np.all(ref[:, 6] == 0)                    # all from ConsToPrim3D
    want = sorted(tuple(int(x) for x in r[:6]) for r in ref)
    assert sorted(rows) == want
    for n, cnt in enumerate(counts):
        for s, c in enumerate(cnt):
            assert c == np.count_nonzero((ref[:, 0] == n) & (ref[:, 1] == s + 1))
