#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the 3D ideal-MHD Godunov step on one MI355X.

Workload (the headline configuration): 256^3 periodic cube, FP64, gamma 5/3,
PLM (MC limiter) + HLLD + RK2 + upwind CT (UCT_HLL) on staggered faces,
CFL 0.4, seeded random-field IC (gpluto_dev_amd.ic.mhd_turbulence).  One
"step" = one full RK2 time step (AdvanceStep) of the whole box; a zone-update
is one cell advanced by one step.

    python bench.py [--gpus 1] [--steps 20] [--warmup 2] [--n 256]

Prints ONE JSON line.  value = FP64 zone-updates/s with the state resident in
HBM (host wall clock around exactly --steps steps replayed from the library's
captured HIP graph, the library stream synchronised on both sides).  The
mixed-precision form runs the same steps from the same IC and is reported
beside it; its L1 error against FP64 is taken at equal times, both forms
stepping on the FP64 run's dt sequence (the free-running comparison, where
each form sets its own dt, is kept beside it).  roofline: the whole step
against HBM, zone-updates/s x the algorithmic bytes of one zone-update (each
stage reads Vc, Vs and Uc once and writes them once: 696 B FP64), with the
measured HBM bytes of one step (rocprofv3 FETCH_SIZE + WRITE_SIZE) beside it;
kernel_roofline: the dominant kernel's algorithmic bytes per launch over its
average launch duration (HIP events on the library's stream, in a second run
of the same steps with per-launch timing on).  cpu_baseline: the oracle on all
host cores and the reference's own serial build (oracle/_ref), each on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s, 3D MHD HLLD+CT 256³: FP64 vs mixed at stated L1; % HBM peak"
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6

# Algorithmic bytes per cell of one launch: each kernel's compulsory HBM
# traffic if every element it touches were read once and written once
# (w = bytes of the state type R, which also types the within-stage sums dU
# and C_dt; the conserved state U and the faces are FP64).  A launch covers
# the whole box; the figure is the mean over the two stages of an RK2 step,
# because the kernels' average launch time is too.  Stage 1 writes its state
# to a second buffer and stage 2 reads both, so no stage copies U^n or the
# faces of U^n aside.  FD holds 4 UCT_HLL records per face (et, eb and the
# two clipped signal speeds).
#   k_sweepx_fused : V 7w + face 8 -> FD 4w + dU 5w (+ C_dt w, stage 1)
#                    = 17w + 8 | 16w + 8                           -> 16.5w + 8
#   k_sweepy_fused : V 7w + face 8 + dU 5w (+ C_dt w) -> FD 4w + dU 5w (+ C_dt w)
#                    = 23w + 8 | 21w + 8                           -> 22w + 8
#   k_sweepz_fused : V 7w + face 8 + dU 5w + U 40 (+ C_dt w | + U^n 40)
#                    -> FD 4w + U 40
#                    = 17w + 88 | 16w + 128                        -> 16.5w + 108
#   k_sweepxy_fused (mixed form: the x and y sweeps in one row march):
#                    V 8w + faces 16 -> FD 8w + dU 5w (+ C_dt w, stage 1)
#                    = 22w + 16 | 21w + 16                         -> 21.5w + 16
#   k_emf          : FD 12w + faces 24 -> E 3w                     =  15w + 24
#   k_finish       : E 3w + faces 24 + U 40 (+ faces of U^n 24, stage 2)
#                    -> faces 24 + V 8w                            =  11w + 100
#   k_sweep_* (split, rows > 512): V 8w + face 8 -> F 5w + CM w + FD 4w = 18w + 8
#   k_rhs_update   : F 15w + CM 3w + U 40 (+ U^n 40, stage 2) -> U 40 = 18w + 100
ALG_BYTES = {"k_sweepx_fused": (16.5, 8), "k_sweepy_fused": (22, 8), "k_sweepxy_fused": (21.5, 16),
             "k_sweepz_fused": (16.5, 108), "k_emf": (15, 24), "k_finish": (11, 100),
             "k_rhs_update": (18, 100)}


def alg_bytes_per_cell(kernel: str, w: int) -> float:
    a, b = ALG_BYTES.get(kernel, (18, 8) if kernel.startswith("k_sweep") else (0, 0))
    return a * w + b


# B_alg of the whole RK2 step per zone-update (4 NV w_V + 5 NF w_B + 5 NU w_U):
B_ALG = {"fp64": 696.0, "mixed": 568.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(n, steps, warmup, precision, vc, vs, first_dt):
    """Warm-up, then exactly ``steps`` steps timed on the host clock as the
    library runs them (graph replay, dt on the device); the final state; then
    the same number of steps again with HIP events around every kernel on the
    library's stream (eager launches) for the per-kernel times."""
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    g = StageUpdate(Definitions(), RuntimeSetup(n=(n, n, n), first_dt=first_dt), precision=precision)
    g.upload(vc, vs)
    err, t, dt = g.run(warmup, first_dt)
    if err:
        raise RuntimeError(f"warmup error bits {err:#x}")
    t0 = time.perf_counter()
    err, t1, dt1 = g.run(steps, -1.0)       # synchronous: returns after the stream drained
    wall = time.perf_counter() - t0
    if err:
        raise RuntimeError(f"error bits {err:#x}")
    vcf, vsf, _ = g.download()
    g.set_timing(True)
    err, _, _ = g.run(steps, -1.0)
    if err:
        raise RuntimeError(f"error bits {err:#x}")
    stats = g.kernel_stats()
    g.close()
    return wall, stats, vcf, vsf, t1


def dominant(stats, w, ncell):
    """Kernel with the largest total time; its bandwidth from algorithmic bytes."""
    best = max(stats.items(), key=lambda kv: kv[1][1])
    name, (cnt, tot) = best
    avg_ms = tot / cnt
    abytes = alg_bytes_per_cell(name, w) * ncell
    return name, avg_ms, abytes


def _traffic_table(w: int):
    path = os.path.join(ROOT, "profiles", "traffic_mhd3d_256.json")
    try:
        with open(path) as f:
            return json.load(f).get("fp64" if w == 8 else "mixed", {})
    except (OSError, ValueError):
        return {}


def traffic_from_profile(name: str, w: int):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    return _traffic_table(w).get(name)


def step_traffic_from_profile(w: int, stages: int = 2):
    """HBM bytes of one whole step (every kernel, once per RK2 stage) from the
    committed rocprofv3 --pmc summary of the 256^3 workload, if any."""
    t = _traffic_table(w)
    return stages * sum(t.values()) if t else None


def cpu_reference_serial(seed: int):
    """The reference's own serial CPU build on a bounded sample (64^3, 20
    steps of the same IC generator); None where the build is absent."""
    from gpluto_dev_amd import ic
    m, steps = 64, 20
    binary = os.path.join(ROOT, "oracle", "_ref", "mhd3d_RK2_MC_LIM_UCT_HLL", "pluto")
    if not os.path.exists(binary):
        return None
    import ref_run
    vc, vs = ic.mhd_turbulence((m, m, m), seed)
    r = ref_run.run_reference(binary, vc, vs, gamma=5 / 3, steps=steps, cfl=0.4,
                              first_dt=5e-4 * 256 / m, solver="hlld")
    mt = re.search(r"Elapsed time \(compute\)\s+([0-9.eE+-]+)", r.log)
    secs = float(mt.group(1)) if mt else r.wall_s
    return {"value": m ** 3 * steps / secs, "unit": "zone-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"{m}^3 box of the same seeded IC generator, {steps} RK2 steps, "
                      f"reference serial build (g++ -O3, oracle/_ref), compute time "
                      f"{secs:.2f} s from its own 'Elapsed time (compute)' line"}


def cpu_baseline(seed: int):
    """The CPU path on all host cores: the oracle (the reference's algorithm
    restated in C, pinned bit for bit to the reference build) with OpenMP over
    pencils and planes, on a bounded sample (128^3, 10 steps of the same IC
    generator); the reference's own serial build beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from gpluto_dev_amd import ic
    import oracle
    m, steps = 128, 10
    vc, vs = ic.mhd_turbulence((m, m, m), seed)
    o = oracle.Oracle(phys="mhd", n=(m, m, m), gamma=5 / 3)
    o.set_state(vc, vs)
    t0 = time.perf_counter()
    o.run(steps, 5e-4 * 256 / m)
    secs = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"value": m ** 3 * steps / secs, "unit": "zone-updates/s", "cores": cores, "kind": "port",
           "sample": f"{m}^3 box of the same seeded IC generator, {steps} RK2 steps, C oracle "
                     f"(OpenMP, {cores} threads), {secs:.2f} s wall"}
    ref = cpu_reference_serial(seed)
    if ref is not None:
        out["reference_serial"] = ref
    return out


def pmc_probe(n, seed, precision):
    """One step for the byte counters, then a plain copy of known size."""
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    vc, vs = ic.mhd_turbulence((n, n, n), seed)
    g = StageUpdate(Definitions(), RuntimeSetup(n=(n, n, n)), precision=precision)
    g.upload(vc, vs)
    err, _, _ = g.run(1, 5e-4 * 256 / n)
    g.close()
    import ctypes
    probe = ctypes.CDLL(os.path.join(ROOT, "gpluto_dev_amd", "libgpl_probe.so"))
    probe.gpl_probe_copy.argtypes = [ctypes.c_size_t]
    probe.gpl_probe_copy.restype = ctypes.c_int
    copy_bytes = 1 << 30
    rc = probe.gpl_probe_copy(copy_bytes)
    if rc:
        raise RuntimeError(f"copy probe failed: hipError {rc}")
    print(json.dumps({"pmc_probe": precision, "n": n, "err": err,
                      "copy_kernel": "k_copy_probe", "copy_read_bytes": copy_bytes,
                      "copy_write_bytes": copy_bytes}), flush=True)


def same_dt_l1(n: int, steps: int, vc, vs, first_dt: float):
    """Per-variable relative L1 of the mixed form against FP64 at the same
    times: both advance with gpl_advance_step on the FP64 run's own dt
    sequence (NextTimeStep of the FP64 run after every step)."""
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    ctx = {}
    for prec in ("fp64", "mixed"):
        ctx[prec] = StageUpdate(Definitions(), RuntimeSetup(n=(n, n, n), first_dt=first_dt), precision=prec)
        ctx[prec].upload(vc, vs)
    dt = first_dt
    for _ in range(steps):
        e64, inv = ctx["fp64"].advance_step(dt)
        em, _ = ctx["mixed"].advance_step(dt)
        if e64 or em:
            raise RuntimeError(f"error bits {e64:#x} / {em:#x}")
        dt = ctx["fp64"].next_time_step(dt, inv)
    (v64, s64, _), (vm, sm, _) = ctx["fp64"].download(), ctx["mixed"].download()
    for g in ctx.values():
        g.close()
    names = ["rho", "vx1", "vx2", "vx3", "Bx1", "Bx2", "Bx3", "prs"]
    l1 = {names[v]: float(np.sum(np.abs(vm[v] - v64[v])) / np.sum(np.abs(v64[v]))) for v in range(8)}
    for d in range(3):
        l1[f"Bx{d+1}s"] = float(np.sum(np.abs(sm[d] - s64[d])) / np.sum(np.abs(s64[d])))
    return l1, ic.divb_max(sm)


def _rmhd_states(m: int, seed: int, steps: int = 10):
    """(conserved, primitive) pairs [m^3, 8] of an RMHD run of the 192^3
    workload's kind at m^3: its final Uc (with the cell field) and Vc."""
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    sigma = float(10.0 ** np.random.Generator(np.random.PCG64(seed + 1000)).uniform(-1.0, 1.0))
    vc, vs = ic.rmhd_turbulence((m, m, m), seed, lorentz_max=10.0, sigma=sigma)
    g = StageUpdate(Definitions(physics="RMHD"),
                    RuntimeSetup(n=(m, m, m), solver="hll", gamma=4.0 / 3.0, cfl=0.4))
    g.upload(vc, vs)
    g.run(steps, 1e-4)
    v, _, u = g.download()
    g.close()
    prim = np.ascontiguousarray(v.reshape(8, -1).T)
    cons = np.ascontiguousarray(u.reshape(8, -1).T)
    cons[:, 4:7] = prim[:, 4:7]
    return cons, prim, sigma


def c2p_probe(mode: str, seed: int):
    """The con2prim solve for counters: 'train' trains the network on a 64^3
    prior run (seed + 1), takes the final states of a 128^3 run (seed) and
    saves both to the temporary directory; 'solve' loads them and runs the
    batched solve from the A27 and from the learned start, twice each."""
    import tempfile
    from gpluto_dev_amd import c2p_net
    path = os.path.join(tempfile.gettempdir(), "gpl_c2p_probe.npz")
    if mode == "train":
        cons, prim, _ = _rmhd_states(64, seed + 1)
        x, y = c2p_net.pairs(cons, prim)
        w, loss = c2p_net.fit(x, y, steps=3000, batch=4096, lr=3e-3, seed=seed)
        u, _, _ = _rmhd_states(128, seed)
        np.savez(path, w=w, u=u, loss=loss)
    else:
        with np.load(path) as z:
            w, u, loss = z["w"], z["u"], float(z["loss"])
    for _ in range(2):
        _, pa, _, ma = c2p_net.solve(u)
        _, pn, _, mn, rs = c2p_net.solve(u, w, restarts=True)

    def wave_max(p):
        return float(p[:p.size // 64 * 64].reshape(-1, 64).max(1).mean())

    print(json.dumps({"c2p_probe": mode, "states": int(u.shape[0]), "loss": loss,
                      "passes": [float(pa.mean()), float(pn.mean())],
                      "wave_max_passes": [wave_max(pa), wave_max(pn)], "restart_rate": float(rs.mean()),
                      "kernel_ms": [ma, mn]}), flush=True)


def run_batch2d(seed: int, steps: int = 100):
    """The batched 2D workload: 64 periodic boxes, N log-uniform in [32, 512),
    beta ~ U[1, 100], PLM(MC) + HLLD + UCT_HLL + RK2, CFL 0.4, per-box dt,
    FP64 and FP32; zone-updates/s over all boxes (autotuned launch shape)."""
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.batch import BatchUpdate
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup
    sizes, beta = ic.mhd_batch_boxes(seed)
    ics = [ic.mhd_turbulence((n, n, 1), ic.box_seed(seed, b), beta=beta[b])
           for b, n in enumerate(sizes)]
    cells = sum(n * n for n in sizes)
    res = {"boxes": len(sizes), "cells": cells, "steps": steps,
           "sizes_min_max": [min(sizes), max(sizes)]}
    finals = {}
    for prec in ("fp64", "mixed"):
        g = BatchUpdate(Definitions(dimensions=2), RuntimeSetup(solver="hlld", cfl=0.4), sizes,
                        precision=prec)
        for b, (vc, vs) in enumerate(ics):
            g.upload(b, vc, vs)
        err, _, _ = g.run(steps, 5e-4)          # includes the autotuner's first steps
        if err:
            raise RuntimeError(f"batch2d {prec}: error bits {err:#x}")
        for b, (vc, vs) in enumerate(ics):      # timed run from the same IC, shape fixed
            g.upload(b, vc, vs)
        t0 = time.perf_counter()
        err, _, _ = g.run(steps, 5e-4)
        wall = time.perf_counter() - t0
        march, bpc, _ = g.tuning()
        finals[prec] = [g.download(b)[0] for b in range(len(sizes))]
        g.close()
        res[prec] = {"value": cells * steps / wall, "unit": "zone-updates/s",
                     "ms_per_step": wall / steps * 1e3, "march": march, "blocks_per_cu": bpc}
    l1 = max(float(np.sum(np.abs(finals["mixed"][b][v] - finals["fp64"][b][v]))
                   / np.sum(np.abs(finals["fp64"][b][v])))
             for b in range(len(sizes)) for v in range(8))
    res["mixed"]["max_l1_rel_vs_fp64"] = l1
    res["mixed"]["speedup_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    return res


def run_rmhd(seed: int, n: int = 192, steps: int = 10):
    """The 3D RMHD workload: 192^3 periodic, gamma 4/3, PLM(MC) + HLL + UCT_HLL
    + RK2, CFL 0.4, 10 steps, Lorentz factors up to 10 and magnetisation
    sigma log-uniform in [0.1, 10] by seed.  The stage update runs with the
    reference's Eq. (A27) con2prim start and with the learned start: an
    8-64-64-1 BF16 network trained here on the GPU on the (conserved,
    primitive) pairs of a prior run (96^3, the next seed, 10 steps).  Newton
    passes and restart rate of both starts are measured on the final
    conserved states of the plain run."""
    from gpluto_dev_amd import c2p_net, ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

    sigma = float(10.0 ** np.random.Generator(np.random.PCG64(seed + 1000)).uniform(-1.0, 1.0))
    vc, vs = ic.rmhd_turbulence((n, n, n), seed, lorentz_max=10.0, sigma=sigma)
    # the prior run whose states train the network
    m = 96
    t0 = time.perf_counter()
    cons, prim, psigma = _rmhd_states(m, seed + 1, steps)
    x, y = c2p_net.pairs(cons, prim)
    w, loss = c2p_net.fit(x, y, steps=3000, batch=4096, lr=3e-3, seed=seed)
    train_s = time.perf_counter() - t0
    res = {"n": n, "zones": n ** 3, "steps": steps, "sigma": sigma, "lorentz_max": 10.0,
           "train": {"seconds": train_s, "loss": loss, "pairs": int(len(y)), "steps": 3000,
                     "source": f"prior run: {m}^3, seed {seed + 1}, sigma {psigma:.3g}, "
                               f"{steps} steps, final (Uc, Vc)"}}
    finals = {}
    for use in (False, True):
        g = StageUpdate(Definitions(physics="RMHD"),
                        RuntimeSetup(n=(n, n, n), solver="hll", gamma=4.0 / 3.0, cfl=0.4))
        if use:
            g.set_c2p_network(w)
        g.upload(vc, vs)
        g.run(1, 1e-4)                            # warm-up (graph capture)
        g.upload(vc, vs)
        t0 = time.perf_counter()
        err, t, _ = g.run(steps, 1e-4)
        wall = time.perf_counter() - t0
        finals[use] = g.download()
        g.close()
        res["learned" if use else "a27"] = {"value": n ** 3 * steps / wall, "unit": "zone-updates/s",
                                            "ms_per_step": wall / steps * 1e3, "err": err, "time": t}
    v0, v1 = finals[False][0], finals[True][0]
    res["learned"]["max_l1_rel_vs_a27"] = max(float(np.sum(np.abs(v1[q] - v0[q])) / np.sum(np.abs(v0[q])))
                                              for q in range(8))
    u = np.ascontiguousarray(finals[False][2].reshape(8, -1).T)
    _, pa, ea, ms_a = c2p_net.solve(u)
    _, pn, en, ms_n, rs = c2p_net.solve(u, w, restarts=True)
    def wave_max(p):     # a wavefront runs Newton until its slowest lane is done
        return float(p[:p.size // 64 * 64].reshape(-1, 64).max(1).mean())

    res["con2prim"] = {"states": int(u.shape[0]),
                       "a27": {"mean_passes": float(pa.mean()), "wave_max_passes": wave_max(pa),
                               "failed": int(np.count_nonzero(ea)), "kernel_ms": ms_a},
                       "learned": {"mean_passes": float(pn.mean()), "wave_max_passes": wave_max(pn),
                                   "failed": int(np.count_nonzero(en)),
                                   "restart_rate": float(rs.mean()), "kernel_ms": ms_n}}
    return res


def run_mixed512(seed: int, steps: int = 20):
    """The 512^3 mixed-precision configuration: the same random-field MHD IC
    at 512^3, 20 steps in FP64 and in the mixed form; per-variable relative
    L1 of the mixed run against the FP64 run and max |div B|."""
    from gpluto_dev_amd import ic
    n = 512
    first_dt = 5e-4 * 256 / n
    log(f"bench: generating {n}^3 IC (seed {seed})")
    vc, vs = ic.mhd_turbulence((n, n, n), seed)
    res = {"n": n, "zones": n ** 3, "steps": steps}
    out = {}
    for prec in ("fp64", "mixed"):
        log(f"bench: {n}^3 {prec} run")
        wall, stats, vcf, vsf, t = run_gpu(n, steps, 2, prec, vc, vs, first_dt)
        out[prec] = (vcf, vsf)
        res[prec] = {"value": n ** 3 * steps / wall, "ms_per_step": wall / steps * 1e3, "time": t,
                     "divb_max": ic.divb_max(vsf)}
    names = ["rho", "vx1", "vx2", "vx3", "Bx1", "Bx2", "Bx3", "prs"]
    (v64, s64), (vm, sm) = out["fp64"], out["mixed"]
    l1 = {names[q]: float(np.sum(np.abs(vm[q] - v64[q])) / np.sum(np.abs(v64[q]))) for q in range(8)}
    res["mixed"]["l1_rel_vs_fp64"] = l1
    res["mixed"]["max_l1_rel"] = max(l1.values())
    res["mixed"]["speedup_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    return res


def run_dump(seed: int, n: int = 256, steps: int = 40):
    """The cost of one dbl dump (data.NNNN.dbl, dbl.out, restart.out) in the
    256^3 FP64 workload: `steps` steps with the dump after the first half,
    written synchronously (gpl_write_dbl) or overlapped with the second half
    (gpl_write_dbl_async, joined by gpl_wait_dbl at the end), against the
    same steps without a dump (best of two runs each).  Files go to a
    temporary directory."""
    import shutil
    import tempfile
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    vc, vs = ic.mhd_turbulence((n, n, n), seed)
    g = StageUpdate(Definitions(), RuntimeSetup(n=(n, n, n), first_dt=5e-4))
    g.upload(vc, vs)
    g.run(3, 5e-4)                               # graph capture
    tmp = tempfile.mkdtemp(prefix="gpl_dump_")
    try:
        g.write_dbl(tmp, 0)                      # staging buffers, first file
        size = os.path.getsize(os.path.join(tmp, "data.0000.dbl"))
        half = steps // 2

        def timed(mode):
            path = os.path.join(tmp, "data.0001.dbl")
            if os.path.exists(path):
                os.remove(path)                  # no writeback of an earlier dump in the timing
            t0 = time.perf_counter()
            g.run(half, 0.0)
            if mode == "sync":
                g.write_dbl(tmp, 1)
            elif mode == "async":
                g.write_dbl_async(tmp, 1)
            g.run(steps - half, 0.0)
            if mode == "async":
                g.wait_dbl()
            return (time.perf_counter() - t0) * 1e3

        best = {m: min(timed(m) for _ in range(2)) for m in ("none", "sync", "async")}
        base, sync, asyn = best["none"], best["sync"], best["async"]
    finally:
        g.close()
        shutil.rmtree(tmp, ignore_errors=True)
    return {"n": n, "steps": steps, "dump_bytes": size, "no_dump_ms": base, "sync_ms": sync,
            "async_ms": asyn, "sync_exposed_ms": sync - base, "async_exposed_ms": asyn - base}


def run_hd1d(steps: int = 400):
    """The 1D HD shock tube (4096 zones, outflow, gamma 1.4, PLM(MC) + HLLC +
    RK2, CFL 0.8): the reference's CPU-runnable case, timed on the GPU."""
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    n = 4096
    vc = ic.hd_riemann(n, 0)
    g = StageUpdate(Definitions(physics="HD", dimensions=1),
                    RuntimeSetup(n=(n, 1, 1), solver="hllc", cfl=0.8, gamma=1.4,
                                 boundary="outflow"))
    g.upload(vc)
    g.run(steps, 1e-4)                           # warm-up (graph capture)
    g.upload(vc)
    t0 = time.perf_counter()
    err, t, _ = g.run(steps, 1e-4)
    wall = time.perf_counter() - t0
    g.close()
    return {"value": n * steps / wall, "unit": "zone-updates/s", "steps": steps,
            "ms_per_step": wall / steps * 1e3, "err": err, "time": t}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched 2D, 1D HD and 3D RMHD workloads reported beside the headline")
    ap.add_argument("--mixed512", action="store_true",
                    help="also run the 512^3 FP64 vs mixed comparison (about 1.5 min of IC generation)")
    ap.add_argument("--dump", action="store_true",
                    help="also time one dbl dump in the 256^3 run, synchronous and overlapped")
    ap.add_argument("--c2p-probe", choices=["train", "solve"], default=None,
                    help="profiling helper: the RMHD con2prim solve, A27 and learned start")
    ap.add_argument("--pmc-probe", choices=["fp64", "mixed"], default=None,
                    help="run one step of the workload plus a 1 GiB device copy (k_copy_probe, the "
                         "HBM byte-counter calibration) for rocprofv3 --pmc, then exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no halo exchange between GPUs); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from gpluto_dev_amd import ic
    n = a.n
    ncell = n ** 3
    if a.c2p_probe:
        c2p_probe(a.c2p_probe, a.seed)
        return
    if a.pmc_probe:
        pmc_probe(n, a.seed, a.pmc_probe)
        return
    first_dt = 5e-4 * 256 / n
    log(f"bench: generating {n}^3 IC (seed {a.seed})")
    vc, vs = ic.mhd_turbulence((n, n, n), a.seed)

    log("bench: FP64 run")
    wall, stats, vc64, vs64, t64 = run_gpu(n, a.steps, a.warmup, "fp64", vc, vs, first_dt)
    value = ncell * a.steps / wall
    name, avg_ms, abytes = dominant(stats, 8, ncell)
    achieved = abytes / (avg_ms * 1e-3) / 1e9
    traffic = traffic_from_profile(name, 8)
    step_gbs = value * B_ALG["fp64"] / 1e9
    step_traffic = step_traffic_from_profile(8) if n == 256 else None

    out = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D ideal MHD {n}^3 periodic, PLM(MC)+HLLD+UCT_HLL CT+RK2, "
                               f"gamma 5/3, CFL 0.4, seeded random-field IC (beta 10)",
                   "n": n, "zones": ncell, "precision": "fp64", "seed": a.seed},
        "roofline": {"bound": "hbm", "basis": "whole step: zone-updates/s x B_alg",
                     "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": step_gbs / HBM_PEAK_GBS, "traffic": step_traffic,
                     "b_alg_per_zone_update": B_ALG["fp64"], "zone_updates_per_step": ncell,
                     "traffic_per_zone_update": step_traffic / ncell if step_traffic else None},
        "kernel_roofline": {"bound": "hbm", "kernel": name, "achieved": achieved,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                            "traffic": traffic, "alg_bytes_per_launch": abytes,
                            "avg_launch_ms": avg_ms},
        "kernels_ms": {k: (v[1] / v[0] if v[0] else None) for k, v in stats.items()},
    }
    if not a.no_mixed:
        log("bench: mixed run")
        wm, sm, vcm, vsm, tm = run_gpu(n, a.steps, a.warmup, "mixed", vc, vs, first_dt)
        vm = ncell * a.steps / wm
        l1 = {}
        names = ["rho", "vx1", "vx2", "vx3", "Bx1", "Bx2", "Bx3", "prs"]
        for v in range(8):
            l1[names[v]] = float(np.sum(np.abs(vcm[v] - vc64[v])) / np.sum(np.abs(vc64[v])))
        for d in range(3):
            l1[f"Bx{d+1}s"] = float(np.sum(np.abs(vsm[d] - vs64[d])) / np.sum(np.abs(vs64[d])))
        log("bench: mixed against FP64 on one dt sequence")
        l1s, divbs = same_dt_l1(n, a.steps, vc, vs, first_dt)
        out["mixed"] = {"value": vm, "speedup_vs_fp64": vm / value,
                        "l1_rel_vs_fp64": l1s, "max_l1_rel": max(l1s.values()), "same_time": True,
                        "divb_max": divbs,
                        "free_running": {"l1_rel_vs_fp64": l1, "max_l1_rel": max(l1.values()),
                                         "same_time": tm == t64, "divb_max": ic.divb_max(vsm)},
                        "kernels_ms": {k: (v[1] / v[0] if v[0] else None) for k, v in sm.items()}}
    out["divb_max_fp64"] = ic.divb_max(vs64)
    if not a.no_extra:
        log("bench: batched 2D and 1D HD workloads")
        out["batch2d"] = run_batch2d(a.seed)
        out["hd1d"] = run_hd1d()
        log("bench: 3D RMHD workload (A27 and learned con2prim start)")
        out["rmhd"] = run_rmhd(a.seed)
    if a.mixed512:
        out["mixed512"] = run_mixed512(a.seed)
    if a.dump:
        log("bench: dbl dump, synchronous and overlapped")
        out["dump"] = run_dump(a.seed)
    if not a.no_cpu:
        log("bench: CPU baseline")
        try:
            out["cpu_baseline"] = cpu_baseline(a.seed)
        except Exception as e:  # reported, never silently replaced
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
