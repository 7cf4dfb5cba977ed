"""Run the reference CPU build (oracle/_ref/<config>/pluto) on a given IC.

Test infrastructure only: this drives the reference's own executable, built
from its sources by oracle/ref/Makefile, through its normal command line
(``-i pluto.ini -maxsteps N``, Src/cmd_line_opt.cpp:89-111) and reads back
its ``.dbl`` output (Src/write_data.cpp:116-136: each variable's active
zones, i fastest; staggered fields carry one extra face along their
direction, Src/bin_io.cpp:319-324).

Step semantics: ``-maxsteps N`` performs N+1 AdvanceStep calls
(Src/main.cpp:363-371), so a run of S steps passes ``-maxsteps S-1`` and
dumps with ``dn = S-1`` so that only the post-loop output fires
(Src/main.cpp:497-500, CheckForOutput :619-623).
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile
import time
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")


def ref_binary(cfg: str, ts: str = "RK2", lim: str = "MC_LIM", emf: str = "UCT_HLL",
               warn: int = 1) -> str:
    name = f"{cfg}_{ts}_{lim}" + (f"_{emf}" if "mhd" in cfg else "") + (f"_W{warn}" if warn != 1 else "")
    return os.path.join(REF_DIR, name, "pluto")


def build_ref(cfg: str, ts: str = "RK2", lim: str = "MC_LIM", emf: str = "UCT_HLL",
              ref_root: str = "/root/reference", warn: int = 1) -> str:
    """Build one reference configuration (needs the reference sources)."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "ref"), f"CFG={cfg}",
                    f"TS={ts}", f"LIM={lim}", f"EMF={emf}", f"REF={ref_root}", f"WARN={warn}",
                    "-j8"], check=True)
    return ref_binary(cfg, ts, lim, emf, warn)


def write_ic(path: str, vc: np.ndarray, vs: np.ndarray | None, gamma: float, periodic: bool):
    nvar, nz, ny, nx = vc.shape
    nstag = 0 if vs is None else vs.shape[0]
    with open(path, "wb") as f:
        f.write(b"GPLIC01\0")
        f.write(struct.pack("<6i", nx, ny, nz, nvar, nstag, 1 if periodic else 0))
        f.write(struct.pack("<d", gamma))
        f.write(np.ascontiguousarray(vc, dtype="<f8").tobytes())
        if vs is not None:
            f.write(np.ascontiguousarray(vs, dtype="<f8").tobytes())


@dataclass
class RefResult:
    vc: np.ndarray          # [nvar, nz, ny, nx]
    vs: np.ndarray | None   # [dims, nz, ny, nx] lower faces (periodic)
    vs_upper: list | None   # extra upper face slab per direction
    time: float
    wall_s: float
    sec_per_step: float | None
    log: str
    restart: dict | None = None


def _grid_line(name: str, n: int) -> str:
    return f"{name}  0.0  {n}  1.0\n"


def _ini(nx, ny, nz, cfl, cfl_max_var, first_dt, solver, sides, dn):
    ini = ["[Grid]\n\n", _grid_line("X1-grid", nx), _grid_line("X2-grid", ny),
           _grid_line("X3-grid", nz), "\n[Time]\n\n", f"CFL  {cfl!r}\n",
           f"CFL_max_var  {cfl_max_var!r}\n", "tstop  1.e30\n", f"first_dt  {first_dt!r}\n",
           "\n[Solver]\n\n", f"Solver  {solver}\n", "\n[Boundary]\n\n"]
    for d, ax in enumerate(("X1", "X2", "X3")):
        ini.append(f"{ax}-beg  {sides[2*d]}\n{ax}-end  {sides[2*d + 1]}\n")
    ini += ["\n[Static Grid Output]\n\n", "uservar 0\n", f"dbl -1.0 {dn} single_file\n",
            "flt -1.0 -1 single_file\n", "vtk -1.0 -1 single_file\n", "tab -1.0 -1\n",
            "ppm -1.0 -1\n", "png -1.0 -1\n", "log 1000000\n", "analysis -1.0 -1\n",
            "\n[Parameters]\n\n", "DUMMY_PAR  0.0\n"]
    return ini


def _run(binary, tmp, args, timeout):
    env = dict(os.environ, GPLUTO_FTZ_DAZ="0")  # IEEE mode (Src/cpu_fp.hpp:10-17)
    t0 = time.perf_counter()
    p = subprocess.run([binary, "-i", "pluto.ini"] + args, cwd=tmp,
                       env=env, capture_output=True, text=True, timeout=timeout)
    wall = time.perf_counter() - t0
    if p.returncode != 0:
        raise RuntimeError(f"reference run failed ({p.returncode}):\n{p.stdout[-3000:]}\n{p.stderr[-3000:]}")
    return p, wall


def _read_last_dump(tmp, nx, ny, nz):
    with open(os.path.join(tmp, "dbl.out")) as f:
        last = f.read().strip().splitlines()[-1].split()
    nfile, t = int(last[0]), float(last[1])
    names = last[6:]
    raw = np.fromfile(os.path.join(tmp, f"data.{nfile:04d}.dbl"), dtype="<f8")
    ncell = nx * ny * nz
    off = 0
    cells = []
    stag = []
    upper = []
    for name in names:
        if name.endswith("s") and name.startswith("Bx"):
            d = int(name[2]) - 1
            shape = [nz, ny, nx]
            shape[2 - d] += 1
            cnt = shape[0] * shape[1] * shape[2]
            a = raw[off:off + cnt].reshape(shape)
            off += cnt
            sl = [slice(None)] * 3
            sl[2 - d] = slice(0, shape[2 - d] - 1)
            stag.append(np.ascontiguousarray(a[tuple(sl)]))
            su = [slice(None)] * 3
            su[2 - d] = slice(shape[2 - d] - 1, shape[2 - d])
            upper.append(np.ascontiguousarray(a[tuple(su)]))
        else:
            cells.append(raw[off:off + ncell].reshape(nz, ny, nx))
            off += ncell
    assert off == raw.size, (off, raw.size)
    return np.stack(cells), (np.stack(stag) if stag else None), (upper if stag else None), t


# ErrorShow's names (Src/debug_tools.cpp:221-238) and the bits they print
# (Src/gpluto.hpp:247-265)
ERR_BITS = {"ERR_NEGATIVE_RHO": 1 << 0, "ERR_NEGATIVE_ENG": 1 << 1, "ERR_NEGATIVE_PRS": 1 << 2,
            "ERR_NEGATIVE_ENTR": 1 << 3, "ERR_CONS2PRIM_FAIL": 1 << 4, "ERR_PRESSURE_FIX": 1 << 5,
            "ERR_ENG_SOLVE_INACCURATE": 1 << 6, "ERR_ENG_SOLVE_NO_SOLUTION": 1 << 7,
            "ERR_SUPERLUMINAL_SPEED": 1 << 8, "ERR_TOO_MANY_ITERATIONS": 1 << 9,
            "ERR_RIEMANN_SOLVE_FAIL": 1 << 10, "ERR_NAN": 1 << 11, "ERR_POINT_VALUE_FAIL": 1 << 12,
            "ERR_IMEX_GUESS_FAIL": 1 << 13, "ERR_QUIT": 1 << 15,
            "ERR_SUPERLUMINAL_RADIATION_FLUX": 1 << 16, "ERR_GAUSSIAN_SOLVE": 1 << 17,
            "ERR_NEGATIVE_ENR": 1 << 18}


def parse_zone_reports(log: str, nghost: int = 2) -> np.ndarray:
    """The zone-by-zone reports of a WARNING_MESSAGES >= 3 build (Where,
    Src/debug_tools.cpp:633-682) as rows (step, stage, i, j, k, bits, func):
    active-zone indices (the log prints Data indices, ghosts included; an
    absent direction prints nothing and gets 0), g_stepNumber, g_intStage,
    the ERR_* bits, and func 0 for ConsToPrim3D, 1 for a Riemann solver."""
    import re
    rows = []
    lines = log.splitlines()
    for q, line in enumerate(lines):
        m = re.match(r"! (\S+)\(\): zone: (.*)", line)
        if not m:
            continue
        func = 0 if m.group(1) == "ConsToPrim3D" else 1
        z = (re.match(r"i = (\d+),", m.group(2)) or re.match(r"\(i,j\) = \((\d+), (\d+)\)", m.group(2))
             or re.match(r"\(i,j,k\) = \((\d+), (\d+), (\d+)\)", m.group(2)))
        idx = [int(v) - nghost for v in z.groups()]
        ijk = idx + [0] * (3 - len(idx))
        st = re.search(r"step: (\d+); t = \S+; g_intStage = (\d+)", lines[q + 1])
        names = re.search(r"error\(s\) found: (\S*)", lines[q + 2]).group(1)
        bits = 0
        for nm in names.split("/"):
            if nm:
                bits |= ERR_BITS[nm]
        rows.append((int(st.group(1)), int(st.group(2)), *ijk, bits, func))
    return np.array(rows, dtype=np.int64).reshape(-1, 7)


def _sec_per_step(log):
    for line in log.splitlines():
        if "Average time/step" in line:
            try:
                return float(line.split()[-2 if line.strip().endswith("(sec)") else -1])
            except ValueError:
                pass
    return None


def run_reference(binary: str, vc: np.ndarray, vs: np.ndarray | None, *, gamma: float,
                  steps: int, cfl: float, first_dt: float, solver: str,
                  periodic: bool = True, cfl_max_var: float = 1.1, sides=None,
                  timeout: float = 3600.0, workdir: str | None = None,
                  restart_steps: int = 0) -> RefResult:
    """Run ``steps`` steps; with ``restart_steps`` > 0, restart the reference
    from its post-loop dump (``-restart 1``, Src/main.cpp:277-283) and run
    that many more steps.  After the first leg the step counter reads
    steps - 1 (Src/main.cpp:336, 482), which RestartGet restores
    (Src/restart.cpp:392, main.cpp:293), so the second leg passes
    ``-maxsteps steps + restart_steps - 2`` and dumps with that dn.  The
    result then holds the second leg's dump, and ``restart`` the first leg's
    files (dbl.out text, restart.out and data.0001.dbl bytes) and state."""
    if steps < 2:
        raise ValueError("the reference dumps its post-loop output only for >= 2 steps")
    nvar, nz, ny, nx = vc.shape
    bnd = "periodic" if periodic else "outflow"
    sides = list(sides) if sides else [bnd] * 6
    own = workdir is None
    tmp = tempfile.mkdtemp(prefix="gpl_ref_") if own else workdir
    leg1 = None
    try:
        with open(os.path.join(tmp, "pluto.ini"), "w") as f:
            f.writelines(_ini(nx, ny, nz, cfl, cfl_max_var, first_dt, solver, sides, steps - 1))
        write_ic(os.path.join(tmp, "gpl_ic.bin"), vc, vs, gamma, periodic)
        p, wall = _run(binary, tmp, ["-maxsteps", str(steps - 1)], timeout)
        if restart_steps > 0:
            rvc, rvs, rup, rt = _read_last_dump(tmp, nx, ny, nz)
            files = {}
            for name in ("dbl.out", "restart.out", "data.0001.dbl"):
                with open(os.path.join(tmp, name), "rb") as f:
                    files[name] = f.read()
            leg1 = dict(vc=rvc, vs=rvs, vs_upper=rup, time=rt, files=files)
            last = steps + restart_steps - 2
            with open(os.path.join(tmp, "pluto.ini"), "w") as f:
                f.writelines(_ini(nx, ny, nz, cfl, cfl_max_var, first_dt, solver, sides, last))
            p, wall = _run(binary, tmp, ["-restart", "1", "-maxsteps", str(last)], timeout)
            for name in ("dbl.out", "restart.out"):
                with open(os.path.join(tmp, name), "rb") as f:
                    leg1["files_final_" + name] = f.read()
        ovc, ovs, oup, t = _read_last_dump(tmp, nx, ny, nz)
    finally:
        if own:
            import shutil
            shutil.rmtree(tmp, ignore_errors=True)
    return RefResult(vc=ovc, vs=ovs, vs_upper=oup, time=t, wall_s=wall,
                     sec_per_step=_sec_per_step(p.stdout), log=p.stdout, restart=leg1)
