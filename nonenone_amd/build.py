"""Build libnonenone_amd.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m nonenone_amd.build          # or __graft_entry__.build()

The kernels for each component count live in their own translation unit
(csrc/nn_nc<NC>.hip) and compile in parallel; csrc/nn_api.hip holds the C ABI.

Flags: -ffp-contract=off keeps +,-,*,/ rounding identical to the reference's
host build (no fused multiply-add), so the flash follows the same iterates.
-Xclang -target-feature -Xclang -packed-fp32-ops is required on these
machines (packed FP32 instructions are disabled).  -O2 rather than -O3: at -O3
the fully inlined flash kernel returned results that depended on which other
states shared the wavefront (a single state, or 64 copies of it, flashed
correctly), while host builds of the same source under ASan/UBSan were clean.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = Path(os.environ.get("NN_LIB_OUT", str(HERE / "libnonenone_amd.so")))
BUILD_DIR = Path(os.environ.get("NN_BUILD_DIR", str(HERE.parent / "_build")))
# Minimum wavefronts per SIMD of the per-state kernels (csrc/launch.hpp).  Only 1
# is built: a build for 2 waves per SIMD (256-VGPR budget, 15.8 KB of private
# segment per lane in flashKernel<3>) stopped on the GPU with a memory-aperture
# violation in its first 262,144-state flash batch, cause not identified.
WAVES_PER_EU = os.environ.get("NN_WAVES_PER_EU")
if WAVES_PER_EU not in (None, "", "1"):
    raise RuntimeError("NN_WAVES_PER_EU: only 1 wavefront per SIMD is built (the 2-wave build "
                       "faulted on the GPU with a memory-aperture violation)")
# Derivative columns per chunk of the assembly kernel's AD (csrc/launch.hpp).
ASSEMBLY_CHUNK = os.environ.get("NN_ASSEMBLY_CHUNK")
# Rachford-Rice recovery work after which a lane hands its state to the
# wavefront-per-state pass (csrc/common.hpp).
DEFER_BUDGET = os.environ.get("NN_DEFER_BUDGET")
# 0: the lane pass runs the SW allocation recovery itself instead of handing the
# state on (csrc/common.hpp).
DEFER_ALLOCATION = os.environ.get("NN_DEFER_ALLOCATION")
ARCH = os.environ.get("NN_OFFLOAD_ARCH", "gfx950")
UNITS = ("nn_api.hip", "nn_nc2.hip", "nn_nc3.hip", "nn_nc4.hip", "nn_nc5.hip", "nn_nc6.hip", "nn_nc8.hip",
         "mlp.hip", "serve.hip")
FLAGS = ["-std=c++17", "-O2", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required)")


def _filtered(stderr: str) -> str:
    # hipcc's host pass warns about the device-only target feature; that is harmless.
    return "\n".join(line for line in stderr.splitlines()
                     if "packed-fp32-ops" not in line and line.strip())


def _compile(unit: str) -> tuple[str, Path, str]:
    obj = BUILD_DIR / (unit + ".o")
    extra = [f"-DNN_WAVES_PER_EU={int(WAVES_PER_EU)}"] if WAVES_PER_EU else []
    if ASSEMBLY_CHUNK:
        extra.append(f"-DNN_ASSEMBLY_CHUNK={int(ASSEMBLY_CHUNK)}")
    if DEFER_BUDGET:
        extra.append(f"-DNN_DEFER_BUDGET={int(DEFER_BUDGET)}")
    if DEFER_ALLOCATION:
        extra.append(f"-DNN_DEFER_ALLOCATION={int(DEFER_ALLOCATION)}")
    cmd = [hipcc(), *FLAGS, *extra, "-Rpass-analysis=kernel-resource-usage",
           "-I", str(HERE.parent / "include"), "-c", str(CSRC / unit), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {unit} ({proc.returncode}):\n{_filtered(proc.stderr)}")
    # per-kernel registers, scratch and occupancy, kept beside the object
    (BUILD_DIR / (unit + ".resources.txt")).write_text(proc.stderr)
    # A kernel whose stack the compiler cannot bound (recursion, indirect calls)
    # would run with a runtime-default stack and overflow its scratch: refuse it.
    kernel = None
    for line in proc.stderr.splitlines():
        if "Function Name:" in line:
            kernel = line.split("Function Name:")[1].split("[")[0].strip()
        if "Dynamic Stack: True" in line:
            raise RuntimeError(f"{unit}: kernel {kernel} has a dynamic (unbounded) stack")
    warnings = "\n".join(line for line in _filtered(proc.stderr).splitlines()
                         if "warning:" in line)
    return unit, obj, warnings


def _up_to_date(unit: str) -> bool:
    """True when the unit's object is newer than the unit and every header it may
    include (csrc/*.hpp, include/*.h*)."""
    obj = BUILD_DIR / (unit + ".o")
    if not obj.exists():
        return False
    deps = [CSRC / unit, *CSRC.glob("*.hpp"), *(HERE.parent / "include").glob("*.h*")]
    return all(d.stat().st_mtime < obj.stat().st_mtime for d in deps)


def build(verbose: bool = True, incremental: bool = False) -> Path:
    """Compile every unit and link.  incremental=True (or NN_BUILD_INCREMENTAL=1)
    reuses objects that are newer than their sources and every header."""
    BUILD_DIR.mkdir(exist_ok=True)
    t0 = time.time()
    incremental = incremental or os.environ.get("NN_BUILD_INCREMENTAL") == "1"
    todo = [u for u in UNITS if not (incremental and _up_to_date(u))]
    with ThreadPoolExecutor(max_workers=max(1, len(todo))) as pool:
        done = {u: r for u, r in zip(todo, pool.map(_compile, todo))}
    results = [done.get(u, (u, BUILD_DIR / (u + ".o"), "")) for u in UNITS]
    tmp = OUT.with_suffix(".so.tmp")
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC",
            *(str(o) for _, o, _ in results), "-o", str(tmp)]
    proc = subprocess.run(link, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{_filtered(proc.stderr)}")
    os.replace(tmp, OUT)
    if verbose:
        for unit, _, warn in results:
            if warn:
                print(f"[{unit}]\n{warn}", file=sys.stderr)
        print(f"built {OUT} in {time.time() - t0:.1f}s", file=sys.stderr)
    return OUT


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
