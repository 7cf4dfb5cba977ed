"""GPU parity of the 3D MHD stage update against the oracle (bitwise, FP64).

The oracle (oracle/gpl_oracle.c) is itself pinned bit-for-bit to the
reference CPU build (tests/test_oracle_golden.py).  Inputs are the seeded
random-field MHD boxes of gpluto_dev_amd.ic; every comparison is on the
same arrays in the same process.
"""
import numpy as np
import pytest

from gpluto_dev_amd import ic
from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

pytestmark = pytest.mark.gpu


def _pair(n, seed, steps, first_dt, ts="RK2", lim="MC_LIM", solver="hlld", precision="fp64",
          sides=None, emf="UCT_HLL"):
    import oracle
    vc, vs = ic.mhd_turbulence(n, seed)
    print(f"case n={n} seed={seed} steps={steps} ts={ts} lim={lim} solver={solver} {precision} "
          f"sides={sides} {emf}", flush=True)
    o = oracle.Oracle(phys="mhd", n=n, gamma=5 / 3, solver=solver, limiter=lim,
                      time_stepping=int(ts[-1]), periodic=sides is None, sides=sides, emf=emf)
    o.set_state(vc, vs)
    oerr, ot, odt = o.run(steps, first_dt)
    ovc, ovs, ouc = o.get_state()
    g = StageUpdate(Definitions(time_stepping=ts, limiter=lim, ct_emf_average=emf),
                    RuntimeSetup(n=n, solver=solver, first_dt=first_dt, sides=tuple(sides or ())),
                    precision=precision)
    g.upload(vc, vs)
    gerr, gt, gdt = g.run(steps, first_dt)
    gvc, gvs, guc = g.download()
    g.close()
    return (oerr, ot, odt, ovc, ovs, ouc), (gerr, gt, gdt, gvc, gvs, guc)


@pytest.mark.parametrize("n,seed,steps,first_dt", [
    ((16, 16, 16), 7, 3, 1e-4),
    ((16, 12, 20), 7, 6, 2e-3),     # non-cubic, CFL limiter active
    ((32, 32, 32), 11, 5, 1e-3),
])
def test_fp64_bitwise_rk2_mc(n, seed, steps, first_dt):
    (oerr, ot, odt, ovc, ovs, ouc), (gerr, gt, gdt, gvc, gvs, guc) = _pair(n, seed, steps, first_dt)
    assert gerr == oerr == 0
    for v in range(8):
        d = np.max(np.abs(gvc[v] - ovc[v]))
        assert np.array_equal(gvc[v], ovc[v]), f"Vc var {v} differs: max {d:.3e}"
    assert np.array_equal(gvs, ovs)
    for v in (0, 1, 2, 3, 7):
        assert np.array_equal(guc[v], ouc[v]), f"Uc var {v} differs"
    assert gt == ot and gdt == odt


def test_fp64_bitwise_rk3_vanleer():
    (oerr, ot, odt, ovc, ovs, _), (gerr, gt, gdt, gvc, gvs, _) = _pair(
        (12, 16, 10), 9, 5, 3e-3, ts="RK3", lim="VANLEER_LIM")
    assert gerr == oerr == 0
    assert np.array_equal(gvc, ovc) and np.array_equal(gvs, ovs)
    assert gt == ot and gdt == odt


def test_fp64_bitwise_hll():
    (oerr, ot, odt, ovc, ovs, _), (gerr, gt, gdt, gvc, gvs, _) = _pair(
        (16, 16, 16), 5, 4, 2e-3, solver="hll")
    assert gerr == oerr == 0
    assert np.array_equal(gvc, ovc) and np.array_equal(gvs, ovs)


@pytest.mark.parametrize("path", ["fused", "split", "emf0"])
def test_fp64_bitwise_hllc(path, monkeypatch):
    """MHD HLLC (Src/MHD/hllc.cpp) with UCT_HLL, on the fused sweeps, the
    per-interface split kernels and the one-edge-per-thread EMF kernel."""
    if path == "split":
        monkeypatch.setenv("GPL_SPLIT_KERNELS", "1")
    elif path == "emf0":
        monkeypatch.setenv("GPL_EMF_MARCH", "0")
    for n, ts, lim in (((16, 12, 20), "RK2", "MC_LIM"), ((20, 17, 12), "RK3", "VANLEER_LIM")):
        (oerr, ot, odt, ovc, ovs, _), (gerr, gt, gdt, gvc, gvs, _) = _pair(
            n, 23, 5, 2e-3, ts=ts, lim=lim, solver="hllc")
        assert gerr == oerr == 0
        assert np.array_equal(gvc, ovc) and np.array_equal(gvs, ovs)
        assert gt == ot and gdt == odt


@pytest.mark.parametrize("path", ["fused", "split", "xy1"])
def test_fp64_bitwise_uct_hlld(path, monkeypatch):
    """HLLD with the UCT_HLLD edge EMF (CT_Store_UCTHlld: dL and dR apart, the
    HLLC limit and the HLL fallback), on the fused and split sweeps."""
    if path == "split":
        monkeypatch.setenv("GPL_SPLIT_KERNELS", "1")
    elif path == "xy1":
        monkeypatch.setenv("GPL_FUSED_XY", "1")
    for n, ts, lim in (((16, 12, 20), "RK2", "MC_LIM"), ((20, 17, 12), "RK3", "VANLEER_LIM")):
        (oerr, ot, odt, ovc, ovs, _), (gerr, gt, gdt, gvc, gvs, _) = _pair(
            n, 25, 5, 2e-3, ts=ts, lim=lim, emf="UCT_HLLD")
        assert gerr == oerr == 0
        assert np.array_equal(gvc, ovc) and np.array_equal(gvs, ovs)
        assert gt == ot and gdt == odt


def test_mixed_precision_uct_hlld():
    (_, _, _, ovc, _, _), (gerr, _, _, gvc, gvs, _) = _pair(
        (32, 24, 20), 26, 8, 1e-3, precision="mixed", emf="UCT_HLLD")
    assert gerr == 0
    for v in range(8):
        l1 = np.sum(np.abs(gvc[v] - ovc[v])) / np.sum(np.abs(ovc[v]))
        assert l1 <= 1e-5, (v, l1)
    assert ic.divb_max(gvs) <= 1e-11 * np.max(np.abs(gvs)) * 32


def test_uct_hlld_needs_hlld():
    """The reference's HLL and HLLC solvers quit with UCT_HLLD
    (Src/MHD/hll.cpp:75-76, Src/MHD/hllc.cpp:81-83)."""
    for solver in ("hll", "hllc"):
        with pytest.raises(RuntimeError, match="not supported"):
            StageUpdate(Definitions(ct_emf_average="UCT_HLLD"), RuntimeSetup(n=(16, 16, 16), solver=solver))
    with pytest.raises(RuntimeError, match="not supported"):
        StageUpdate(Definitions(physics="RMHD", ct_emf_average="UCT_HLLD"),
                    RuntimeSetup(n=(16, 16, 16), solver="hll"))


def test_mixed_precision_hllc():
    (_, _, _, ovc, _, _), (gerr, _, _, gvc, gvs, _) = _pair(
        (32, 24, 20), 24, 8, 1e-3, solver="hllc", precision="mixed")
    assert gerr == 0
    for v in range(8):
        l1 = np.sum(np.abs(gvc[v] - ovc[v])) / np.sum(np.abs(ovc[v]))
        assert l1 <= 1e-5, (v, l1)
    assert ic.divb_max(gvs) <= 1e-11 * np.max(np.abs(gvs)) * 32


def test_advance_step_seam_matches_oracle():
    """AdvanceStep + host NextTimeStep, step by step (Src/main.cpp:446-463)."""
    import oracle
    n = (16, 16, 16)
    vc, vs = ic.mhd_turbulence(n, 3)
    o = oracle.Oracle(phys="mhd", n=n, gamma=5 / 3)
    o.set_state(vc, vs)
    g = StageUpdate(Definitions(), RuntimeSetup(n=n))
    g.upload(vc, vs)
    dt = 2e-3
    for step in range(4):
        oerr, oinv = o.advance(dt)
        gerr, ginv = g.advance_step(dt)
        assert gerr == oerr == 0
        assert ginv == oinv, (step, ginv, oinv)
        dt = g.next_time_step(dt, ginv)
    gvc, gvs, _ = g.download()
    ovc, ovs, _ = o.get_state()
    assert np.array_equal(gvc, ovc) and np.array_equal(gvs, ovs)
    g.close()


@pytest.mark.parametrize("env", ["", "GPL_EMF_MARCH=0", "GPL_FUSED_XY=0"])
def test_mixed_precision_l1_within_target(env, monkeypatch):
    """Mixed form: FP32 V/PLM/HLLD, FP64 U, faces, CT and dt.  Relative L1 per
    variable against the FP64 oracle stays <= 1e-5; div B stays at FP64
    round-off because the face update is FP64.  Default kernels, the
    LDS-staged EMF tiles, and separate x and y sweeps."""
    if env:
        monkeypatch.setenv(*env.split("="))
    n, steps = (32, 32, 32), 10
    (oerr, ot, odt, ovc, ovs, _), (gerr, gt, gdt, gvc, gvs, _) = _pair(
        n, 13, steps, 1e-3, precision="mixed")
    assert gerr == oerr == 0
    for v in range(8):
        l1 = np.sum(np.abs(gvc[v] - ovc[v])) / np.sum(np.abs(ovc[v]))
        print(f"var {v}: rel L1 {l1:.3e}")
        assert l1 <= 1e-5, (v, l1)
    for d in range(3):
        l1 = np.sum(np.abs(gvs[d] - ovs[d])) / np.sum(np.abs(ovs[d]))
        assert l1 <= 1e-5, (d, l1)
    assert ic.divb_max(gvs) <= 1e-11 * np.max(np.abs(gvs)) * n[0]


@pytest.mark.parametrize("path", ["split", "march8", "march16", "march32", "xy1",
                                  "emf0", "emf8", "emf-4", "emf-8", "emf-16", "emf-32"])
def test_kernel_paths_bitwise(path, monkeypatch):
    """Split (per-interface) and fused kernels (the LDS x-row kernel and a
    register-window y march, or x and y rows in one marching kernel (xy1); the
    z march; every march length), and every EMF kernel (one edge per thread,
    register march of 8, LDS-staged tiles marching 4 to 32 planes), including
    segments and tiles that do not divide the box (ny = 12, nz = 20, 17, 31,
    nx = 70)."""
    if path == "split":
        monkeypatch.setenv("GPL_SPLIT_KERNELS", "1")
    elif path.startswith("march"):
        monkeypatch.setenv("GPL_MARCH", path[5:])
    elif path.startswith("emf"):
        monkeypatch.setenv("GPL_EMF_MARCH", path[3:])
    elif path == "xy1":
        monkeypatch.setenv("GPL_FUSED_XY", "1")
    for n in ((16, 12, 20), (20, 17, 31), (70, 9, 6)):
        (oerr, ot, odt, ovc, ovs, _), (gerr, gt, gdt, gvc, gvs, _) = _pair(n, 21, 5, 2e-3)
        assert gerr == oerr == 0
        assert np.array_equal(gvc, ovc) and np.array_equal(gvs, ovs)
        assert gt == ot and gdt == odt


OUT = ("outflow",) * 6
XPER = ("periodic", "periodic") + ("outflow",) * 4          # periodic in x, outflow in y, z


@pytest.mark.parametrize("n,sides,ts,lim,solver,steps", [
    ((16, 12, 10), OUT, "RK2", "MC_LIM", "hlld", 8),
    ((12, 14, 16), XPER, "RK3", "VANLEER_LIM", "hlld", 6),
    ((14, 10, 12), OUT, "RK2", "MC_LIM", "hll", 6),
    ((20, 17, 31), ("outflow", "outflow", "periodic", "periodic", "outflow", "outflow"),
     "RK2", "MC_LIM", "hlld", 5),
    ((14, 12, 10), ("reflective",) * 6, "RK2", "MC_LIM", "hlld", 8),
    ((12, 16, 14), ("reflective", "outflow", "periodic", "periodic", "outflow", "reflective"),
     "RK3", "VANLEER_LIM", "hll", 6),
    ((12, 10, 14), OUT, "RK3", "VANLEER_LIM", "hllc", 6),
    ((14, 12, 10), ("reflective", "outflow", "outflow", "reflective", "periodic", "periodic"),
     "RK2", "MC_LIM", "hlld+UCT_HLLD", 6),
    ((10, 14, 12), ("reflective",) * 2 + ("periodic",) * 2 + ("outflow",) * 2,
     "RK2", "MC_LIM", "hllc", 6),
])
def test_outflow_sides_bitwise(n, sides, ts, lim, solver, steps):
    """Outflow sides (OutflowBoundary + CT_FillStaggeredFields every stage) and
    reflective walls (ReflectiveBoundary) on the padded device box, alone and
    mixed with each other and with periodic directions."""
    solver, _, emf = solver.partition("+")
    (oerr, ot, odt, ovc, ovs, ouc), (gerr, gt, gdt, gvc, gvs, guc) = _pair(
        n, 17, steps, 2e-3, ts=ts, lim=lim, solver=solver, sides=sides, emf=emf or "UCT_HLL")
    assert gerr == oerr == 0
    for v in range(8):
        d = np.max(np.abs(gvc[v] - ovc[v]))
        assert np.array_equal(gvc[v], ovc[v]), f"Vc var {v} differs: max {d:.3e}"
    assert np.array_equal(gvs, ovs)
    for v in (0, 1, 2, 3, 7):
        assert np.array_equal(guc[v], ouc[v]), f"Uc var {v} differs"
    assert gt == ot and gdt == odt


def test_outflow_sides_mixed_precision():
    (oerr, _, _, ovc, _, _), (gerr, _, _, gvc, gvs, _) = _pair(
        (24, 20, 16), 18, 8, 1e-3, precision="mixed", sides=OUT)
    assert gerr == oerr == 0
    for v in range(8):
        l1 = np.sum(np.abs(gvc[v] - ovc[v])) / np.sum(np.abs(ovc[v]))
        assert l1 <= 1e-5, (v, l1)
