"""Build recipe for ``libabm_hip.so`` (gfx950), in-tree.

Each ``csrc/*.hip`` file is compiled to an object with hipcc and the objects
are linked into ``abm_amd/libabm_hip.so``.  The exact float64 translation units
are compiled with ``-ffp-contract=off`` so that no multiply-add is fused where
the reference rounds twice; the fused float32 unit allows contraction (its
float64 re-check functions disable it locally with a pragma).  All HIP code
is built without packed-FP32 instructions.

Usage: ``python -m abm_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libabm_hip.so"
OBJ_DIR = PKG / "_build"
ARCH = os.environ.get("ABM_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]
# per translation unit: fp contraction policy
UNITS = {
    "capi.hip": ["-ffp-contract=off"],
    "vf_f64.hip": ["-ffp-contract=off"],
    "vf_f32.hip": ["-ffp-contract=fast"],
    "vf_group.hip": ["-ffp-contract=fast"],
    "forage.hip": ["-ffp-contract=off"],
    "mt19937.hip": [],
    "cs.hip": ["-ffp-contract=off"],
    "stats.hip": ["-ffp-contract=off"],
    "agentlog.hip": [],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build(verbose: bool = False) -> Path:
    """Compile every unit and link the shared library; returns its path."""
    cc = hipcc()
    OBJ_DIR.mkdir(exist_ok=True)
    procs = []
    objs = []
    for name, extra in UNITS.items():
        src = CSRC / name
        obj = OBJ_DIR / (src.stem + ".o")
        cmd = [cc, *COMMON, *extra, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((name, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for name, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            failed.append(name)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
