"""Build the HIP library (gfx950) in-tree: gpluto_dev_amd/libgpluto_amd.so.

    python -m gpluto_dev_amd.build        # or: python gpluto_dev_amd/build.py

Flags that matter for results:
  -ffp-contract=off   no fused multiply-add: the FP64 drop-in keeps the
                      reference's rounded multiply + rounded add (its x86-64
                      g++ build emits no FMA), so results match bit for bit.
  -target-feature -packed-fp32-ops
                      no packed FP32 VALU ops (v_pk_*_f32) anywhere.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgpluto_amd.so")
# the byte-counter calibration copy used when profiling (not part of the ABI)
PROBE_LIB = os.path.join(HERE, "libgpl_probe.so")
ARCH = os.environ.get("GPL_OFFLOAD_ARCH", "gfx950")

# (source, extra flags): the FP64 drop-in keeps IEEE division/sqrt and no
# contraction; the mixed form's FP32 kernels may contract, use the fast FP32
# divide/sqrt and flush FP32 denormals, which drops the operand scaling
# around every FP32 reciprocal (its accuracy is stated against the FP64 run;
# FP64 arithmetic in the same files keeps denormals).
MIXED = ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
         "-fgpu-flush-denormals-to-zero"]
# the learned con2prim guess's matrix-core code (c2p_net.hpp): MFMA results
# in ordinary VGPRs, which the ReLU and BF16 conversion read directly, instead
# of accumulation registers copied out one v_accvgpr_read per float
MFMA_VGPR = ["-mllvm", "-amdgpu-mfma-vgpr-form"]
SOURCES = [("mhd3d_fp64.hip", MFMA_VGPR),
           ("mhd3d_mixed.hip", MIXED),
           ("mhd3d_common.hip", []),
           ("diagnostics.hip", []),
           ("mhd3d_boundary.hip", []),
           ("dump_gather.hip", []),
           ("hd_kernels.hip", []),
           ("batch2d_fp64.hip", []),
           ("batch2d_mixed.hip", MIXED),
           ("batch2d_common.hip", []),
           ("batch2d_boundary.hip", []),
           ("engine_batch2d.cpp", []),
           ("c2p_net.hip", MFMA_VGPR),
           ("engine_mhd3d.cpp", []),
           ("engine_hd.cpp", []),
           ("dbl_io.cpp", []),
           ("gpl_api.cpp", [])]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I" + os.path.join(HERE, "..", "include")]


def _link(cc: str, objs: list[str], out: str) -> None:
    tmp = out + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"link failed: {out}")
    os.replace(tmp, out)


def _compile(cc: str, objdir: str, s: str, extra: list[str], verbose: bool) -> str:
    src = os.path.join(CSRC, s)
    obj = os.path.join(objdir, s + ".o")
    cmd = [cc, *flags(), *extra, "-c", src, "-o", obj]
    if s.endswith(".cpp"):
        cmd.insert(1, "-xhip")
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {s}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    units = SOURCES + [("copy_probe.hip", [])]
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _compile(cc, objdir, u[0], u[1], verbose), units))
    _link(cc, objs[:-1], LIB)
    _link(cc, objs[-1:], PROBE_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
