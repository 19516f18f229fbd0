"""Build the gfx950 HIP library in-tree: ``python -m gigalens_amd.build``.

The library is several translation units compiled in parallel (the fused
particle kernels by profile combination, the staged pipeline, the network, the
batched SVI and HMC kernels, the noise sampler and the C ABI) and linked into
``_lib/libgigalens_amd.so``.  ``build()`` compiles
every unit each time it is called; ``build(incremental=True)`` recompiles only
units whose recorded dependencies (``-MMD`` files) changed.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "_lib", "libgigalens_amd.so")
UNITS = ["fused_0.hip", "fused_1.hip", "fused_2.hip", "band.hip", "staged.hip", "cnn.hip", "svi.hip", "hmc.hip", "noise.hip", "capi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for every HIP translation unit.
FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _deps(dfile: str):
    try:
        text = open(dfile).read().replace("\\\n", " ")
    except OSError:
        return None
    _, _, rest = text.partition(":")
    return [p for p in rest.split() if p]


def _stale(src: str, obj: str, dfile: str) -> bool:
    if not os.path.exists(obj):
        return True
    deps = _deps(dfile)
    if not deps:
        return True
    t = os.path.getmtime(obj)
    return any(not os.path.exists(p) or os.path.getmtime(p) > t for p in deps + [src, __file__])


def _compile(unit: str, verbose: bool):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJ, unit.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-MMD", "-MF", obj[:-2] + ".d", "-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    return unit, res.returncode, res.stdout + res.stderr


def build(verbose: bool = False, incremental: bool = False, jobs: int = 0) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    todo = []
    for u in UNITS:
        obj = os.path.join(OBJ, u.replace(".hip", ".o"))
        if not incremental or _stale(os.path.join(CSRC, u), obj, obj[:-2] + ".d"):
            todo.append(u)
    jobs = jobs or min(len(UNITS), max(1, min(8, os.cpu_count() or 1)))
    failed = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for unit, rc, log in ex.map(lambda u: _compile(u, verbose), todo):
            if rc != 0:
                failed.append(unit)
                sys.stderr.write(log)
    if failed:
        raise RuntimeError(f"hipcc failed for {', '.join(failed)}")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + \
        [os.path.join(OBJ, u.replace(".hip", ".o")) for u in UNITS]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"link failed with code {res.returncode}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
