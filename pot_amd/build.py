"""Build recipe of ``libpot_amd.so`` (gfx950 only).

Compiles every ``csrc/*.hip`` with hipcc for ``--offload-arch=gfx950`` in
parallel and links them into ``pot_amd/libpot_amd.so``.  Packed fp32 VALU ops
are switched off for all device code (``-packed-fp32-ops``).

Usage: ``python -m pot_amd.build`` (or ``__graft_entry__.build()``).
"""

import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpot_amd.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(os.path.dirname(HERE), "include"),
]


# Per-file flags.  The online-cost pass and the fused network kernels keep
# their MFMA accumulators in VGPRs, so their epilogues read them without
# accvgpr copies.
EXTRA = {
    "online.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "mlp_fused.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def _sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, src[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj, time.time() - t0


def build(verbose=True, jobs=None):
    """Compile all kernels and link the shared library; returns its path."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for (obj, dt), src in zip(results, srcs):
            print(f"  hipcc {src}: {dt:.1f}s", flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *[o for o, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    shutil.move(tmp, LIB)
    if verbose:
        print(f"  linked {LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
