"""GPU parity of the batched 2D MHD update: every box of a batch of uneven
sizes computes, bit for bit (FP64), what a standalone 2D run of that box
computes -- against the oracle (itself pinned bitwise to the reference CPU
build) and against the reference build's own 2D output fixture.  Launch
shapes chosen by the autotuner never change a bit; the FP32 variant stays
within 1e-5 relative L1 of the FP64 batch.  Both EMF averages: UCT_HLL, and
UCT_HLLD with HLLD (against two more reference fixtures).  Periodic boxes, and
boxes with outflow and reflective sides, alone and mixed with periodic
directions (the padded boxes, against the oracle and two more reference
fixtures, the evolved upper boundary faces included).  A batch runs its boxes
as two groups on two streams; one stream (GPL_BATCH_STREAMS=1) gives the same
bits."""
import numpy as np
import pytest

from golden_io import load
from gpluto_dev_amd import ic
from gpluto_dev_amd.batch import BatchUpdate
from gpluto_dev_amd.stage import Definitions, RuntimeSetup

pytestmark = pytest.mark.gpu


def _batch(sizes, ts="RK2", lim="MC_LIM", solver="hlld", precision="fp64", emf="UCT_HLL", sides=None):
    return BatchUpdate(Definitions(dimensions=2, time_stepping=ts, limiter=lim, ct_emf_average=emf),
                       RuntimeSetup(solver=solver, cfl=0.4, sides=tuple(sides) if sides else ()),
                       sizes, precision=precision)


def _ics(sizes, seed):
    out = []
    for b, (nx, ny) in enumerate(sizes):
        beta = 1.0 + 99.0 * ((b * 37) % 100) / 100.0
        out.append(ic.mhd_turbulence((nx, ny, 1), ic.box_seed(seed, b), beta=beta))
    return out


def _oracle(n, vc, vs, steps, first_dt, ts="RK2", lim="MC_LIM", solver="hlld", emf="UCT_HLL",
            sides=None):
    import oracle
    o = oracle.Oracle(phys="mhd", n=n, gamma=5 / 3, solver=solver, limiter=lim,
                      time_stepping=int(ts[-1]), cfl=0.4, emf=emf, periodic=sides is None, sides=sides)
    o.set_state(vc, vs)
    err, t, dt = o.run(steps, first_dt)
    v, s, u = o.get_state()
    return err, t, dt, v, s, u


def _divb_interior(vs):
    """max |div B| over the cells whose faces are all lower faces of the array
    (a box with walls or outflow sides is not periodic: its upper boundary
    faces are not face 0)."""
    _, _, ny, nx = vs.shape
    d = (vs[0][:, :-1, 1:] - vs[0][:, :-1, :-1]) * nx + (vs[1][:, 1:, :-1] - vs[1][:, :-1, :-1]) * ny
    return float(np.max(np.abs(d)))


SIZES = [(5, 5), (33, 33), (63, 64), (64, 63), (65, 65), (100, 37), (130, 130), (32, 24)]


@pytest.mark.parametrize("ts,lim,solver,emf", [("RK2", "MC_LIM", "hlld", "UCT_HLL"),
                                               ("RK3", "VANLEER_LIM", "hll", "UCT_HLL"),
                                               ("RK2", "VANLEER_LIM", "hllc", "UCT_HLL"),
                                               ("RK3", "MC_LIM", "hlld", "UCT_HLLD")])
def test_every_box_matches_its_standalone_oracle_run(ts, lim, solver, emf):
    steps, first_dt = 6, 1e-3
    print(f"case {ts} {lim} {solver} {emf} sizes={SIZES}", flush=True)
    ics = _ics(SIZES, 4)
    g = _batch(SIZES, ts, lim, solver, emf=emf)
    for b, (vc, vs) in enumerate(ics):
        g.upload(b, vc, vs)
    err, t, dt = g.run(steps, first_dt)
    assert err == 0
    for b, (nx, ny) in enumerate(SIZES):
        vc, vs = ics[b]
        oerr, ot, odt, ov, osv, ou = _oracle((nx, ny, 1), vc, vs, steps, first_dt, ts, lim, solver, emf)
        gv, gs, gu = g.download(b)
        assert oerr == 0
        assert t[b] == ot and dt[b] == odt, f"box {b} ({nx}x{ny}): time/dt differ"
        for v in range(8):
            assert np.array_equal(gv[v], ov[v]), \
                f"box {b} ({nx}x{ny}) var {v}: max diff {np.max(np.abs(gv[v] - ov[v])):.3e}"
        assert np.array_equal(gs, osv), f"box {b}: faces differ"
        for v in (0, 1, 2, 3, 6, 7):
            assert np.array_equal(gu[v], ou[v]), f"box {b}: Uc var {v} differs"
    g.close()


@pytest.mark.parametrize("name", ["mhd2d_32x24_rk2_mc", "mhd2d_30x26_rk2_mc_ucthlld",
                                  "mhd2d_26x34_rk3_vl_ucthlld", "mhd2d_32x24_rk2_mc_outflow",
                                  "mhd2d_28x20_rk2_mc_walls"])
def test_reference_fixture_box_inside_a_batch(name):
    """The reference build's 2D output, reproduced by one box of a batch, the
    upper boundary faces included (evolved on an outflow or reflective side,
    the image of face 0 on a periodic one)."""
    p, a = load(name)
    print(f"case {name}", flush=True)
    nx, ny, _ = p["n"]
    sizes = [(40, 40), (nx, ny), (70, 33)]
    ics = _ics(sizes, 9)
    sides = p.get("sides") or (None if p["periodic"] else ("outflow",) * 6)
    g = _batch(sizes, p["time_stepping"], p["limiter"], p["solver"], emf=p.get("emf", "UCT_HLL"),
               sides=sides)
    for b, (vc, vs) in enumerate(ics):
        if b == 1:
            vc, vs = a["vc_in"], a["vs_in"]
        g.upload(b, vc, vs)
    err, t, _ = g.run(p["steps"], p["first_dt"])
    gv, gs, _ = g.download(1)
    up = g.upper_faces(1)
    g.close()
    assert err == 0
    assert abs(t[1] - p["time"]) <= 1e-6 * abs(p["time"])
    assert np.array_equal(gv, a["vc_out"])
    assert np.array_equal(gs, a["vs_out"])
    for d in range(2):
        assert np.array_equal(up[d], a[f"vs_upper{d}"]), f"upper face {d}"


def test_launch_shapes_and_autotuner_give_identical_bits():
    sizes = [(47, 47), (130, 130), (75, 75), (300, 300), (34, 34), (200, 199)]
    ics = _ics(sizes, 2)
    results, tuned = [], None
    for shape in [None, (8, 2), (32, 0), (16, 4)]:
        print(f"shape {shape}", flush=True)
        g = _batch(sizes)
        if shape is not None:
            g.set_tuning(*shape)
        for b, (vc, vs) in enumerate(ics):
            g.upload(b, vc, vs)
        err, t, dt = g.run(30, 5e-4)
        assert err == 0
        if shape is None:
            tuned = g.tuning()
        results.append((t, dt, [g.download(b) for b in range(len(sizes))]))
        g.close()
    assert tuned[0] in (8, 16, 32) and tuned[1] in (0, 2, 4, 8) and tuned[2] > 0
    t0, dt0, s0 = results[0]
    for t, dt, s in results[1:]:
        assert np.array_equal(t, t0) and np.array_equal(dt, dt0)
        for (a1, b1, c1), (a2, b2, c2) in zip(s, s0):
            assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and np.array_equal(c1, c2)


@pytest.mark.parametrize("emf,sides", [("UCT_HLL", None), ("UCT_HLLD", None),
                                       ("UCT_HLL", ("reflective", "outflow") * 2 + ("outflow",) * 2)])
def test_fp32_batch_within_1e5_of_fp64(emf, sides):
    sizes = [(64, 64), (150, 150), (37, 37), (256, 256)]
    ics = _ics(sizes, 6)
    out = {}
    for prec in ("fp64", "mixed"):
        print(f"case {emf} {sides} {prec}", flush=True)
        g = _batch(sizes, precision=prec, emf=emf, sides=sides)
        for b, (vc, vs) in enumerate(ics):
            g.upload(b, vc, vs)
        err, t, _ = g.run(20, 5e-4)
        assert err == 0
        out[prec] = [g.download(b) for b in range(len(sizes))]
        g.close()
    for b in range(len(sizes)):
        v64, s64, _ = out["fp64"][b]
        v32, s32, _ = out["mixed"][b]
        for v in range(8):
            l1 = np.sum(np.abs(v32[v] - v64[v])) / np.sum(np.abs(v64[v]))
            assert l1 <= 1e-5, f"box {b} var {v}: rel L1 {l1:.2e}"
        assert (ic.divb_max(s32) if sides is None else _divb_interior(s32)) < 1e-10


@pytest.mark.parametrize("precision,emf,sides", [("fp64", "UCT_HLL", None), ("mixed", "UCT_HLLD", None),
                                                 ("mixed", "UCT_HLL", ("outflow", "reflective") * 2 + ("outflow",) * 2)])
def test_two_box_groups_match_one_stream(precision, emf, sides, monkeypatch):
    """The two box groups on two streams (the default), and four groups,
    against the whole batch on one stream: the same bits, times and dts,
    through the autotuned graph replay and the eager steps after it."""
    sizes = [(64, 64), (150, 150), (37, 37), (256, 256), (5, 5), (100, 37)]
    ics = _ics(sizes, 13)
    out = {}
    for streams in ("1", "2", "4"):
        monkeypatch.setenv("GPL_BATCH_STREAMS", streams)
        print(f"case {precision} {emf} {sides} streams={streams}", flush=True)
        g = _batch(sizes, precision=precision, emf=emf, sides=sides)
        for b, (vc, vs) in enumerate(ics):
            g.upload(b, vc, vs)
        err, t, dt = g.run(61, 5e-4)
        assert err == 0
        out[streams] = (t, dt, [g.download(b) for b in range(len(sizes))])
        g.close()
    t1, dt1, s1 = out["1"]
    for streams in ("2", "4"):
        t2, dt2, s2 = out[streams]
        assert np.array_equal(t1, t2) and np.array_equal(dt1, dt2), f"{streams} groups: time/dt differ"
        for b, ((a1, f1, u1), (a2, f2, u2)) in enumerate(zip(s1, s2)):
            assert np.array_equal(a1, a2), f"{streams} groups, box {b} {sizes[b]}: cells differ"
            assert np.array_equal(f1, f2), f"{streams} groups, box {b} {sizes[b]}: faces differ"
            assert np.array_equal(u1, u2), f"{streams} groups, box {b} {sizes[b]}: Uc differs"


def test_uct_hlld_needs_hlld():
    """CT_Store_UCTHlld reads the HLLD star states: the reference's HLL and
    HLLC solvers quit with it (Src/MHD/hll.cpp:75-76, Src/MHD/hllc.cpp:81-83)."""
    for solver in ("hll", "hllc"):
        with pytest.raises(RuntimeError, match="not supported"):
            _batch([(16, 16)], solver=solver, emf="UCT_HLLD")


SIDE_CASES = [
    (("outflow",) * 4, "RK2", "MC_LIM", "hlld", "UCT_HLL"),
    (("reflective",) * 4, "RK3", "VANLEER_LIM", "hllc", "UCT_HLL"),
    (("periodic", "periodic", "outflow", "reflective"), "RK2", "VANLEER_LIM", "hlld", "UCT_HLLD"),
    (("reflective", "outflow", "periodic", "periodic"), "RK2", "MC_LIM", "hll", "UCT_HLL"),
]
SIDE_SIZES = [(5, 5), (33, 20), (64, 63), (40, 70), (100, 37)]


@pytest.mark.parametrize("sides,ts,lim,solver,emf", SIDE_CASES)
def test_sides_every_box_matches_its_standalone_oracle_run(sides, ts, lim, solver, emf):
    """Outflow sides (OutflowBoundary + CT_FillStaggeredFields at every stage)
    and reflective walls on padded boxes, alone and mixed with periodic
    directions: every box bit for bit as its standalone oracle run."""
    steps, first_dt = 8, 1e-3
    sides6 = tuple(sides) + ("outflow", "outflow")
    print(f"case {sides} {ts} {lim} {solver} {emf} sizes={SIDE_SIZES}", flush=True)
    ics = _ics(SIDE_SIZES, 7)
    g = _batch(SIDE_SIZES, ts, lim, solver, emf=emf, sides=sides6)
    for b, (vc, vs) in enumerate(ics):
        g.upload(b, vc, vs)
    err, t, dt = g.run(steps, first_dt)
    assert err == 0
    for b, (nx, ny) in enumerate(SIDE_SIZES):
        vc, vs = ics[b]
        oerr, ot, odt, ov, osv, ou = _oracle((nx, ny, 1), vc, vs, steps, first_dt, ts, lim, solver, emf,
                                             sides=sides6)
        gv, gs, gu = g.download(b)
        assert oerr == 0
        assert t[b] == ot and dt[b] == odt, f"box {b} ({nx}x{ny}): time/dt differ"
        for v in range(8):
            assert np.array_equal(gv[v], ov[v]), \
                f"box {b} ({nx}x{ny}) var {v}: max diff {np.max(np.abs(gv[v] - ov[v])):.3e}"
        assert np.array_equal(gs, osv), f"box {b}: faces differ"
        for v in (0, 1, 2, 3, 6, 7):
            assert np.array_equal(gu[v], ou[v]), f"box {b}: Uc var {v} differs"
    g.close()


def test_one_sided_periodic_direction_is_rejected():
    with pytest.raises(RuntimeError, match="invalid argument"):
        _batch([(16, 16)], sides=("periodic", "outflow") + ("outflow",) * 4)


def test_2d_context_mixed_within_1e5_of_fp64():
    """A 2D MHD gpl_create context (one box of these kernels) in the mixed form,
    with walls and outflow sides, stays within 1e-5 relative L1 of its FP64 run."""
    from gpluto_dev_amd.stage import StageUpdate
    n = (96, 80, 1)
    vc, vs = ic.mhd_turbulence(n, 12)
    sides = ("reflective", "outflow", "outflow", "reflective", "outflow", "outflow")
    out = {}
    for prec in ("fp64", "mixed"):
        print(f"case 2D context {prec}", flush=True)
        g = StageUpdate(Definitions(dimensions=2), RuntimeSetup(n=n, sides=sides), precision=prec)
        g.upload(vc, vs)
        err, _, _ = g.run(20, 5e-4)
        assert err == 0
        out[prec] = g.download()
        g.close()
    (v64, s64, _), (v32, s32, _) = out["fp64"], out["mixed"]
    for v in range(8):
        l1 = np.sum(np.abs(v32[v] - v64[v])) / np.sum(np.abs(v64[v]))
        assert l1 <= 1e-5, f"var {v}: rel L1 {l1:.2e}"
    assert _divb_interior(s32) < 1e-10
