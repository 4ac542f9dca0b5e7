"""Build the gfx950 C-ABI library ``torchwi_amd/lib/libtorchwi_hip.so``.

    python -m torchwi_amd.build

Flags that matter for results:
  -ffp-contract=off                        no fused multiply-add: the per-cell
                                           arithmetic must round exactly as the
                                           reference CPU path does
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE fp32 division (Keys, virtual source)
  -fno-gpu-flush-denormals-to-zero         keep fp32 denormals, as the CPU does
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed fp32 instructions
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "td2d_hip.hip")
DEPS = [SRC, os.path.join(ROOT, "include", "torchwi_hip.h")] + [
    os.path.join(HERE, "csrc", f) for f in ("td2d_arith.hpp", "td2d_vec.hpp", "td2d_pipe.hpp")]
OUT_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(OUT_DIR, "libtorchwi_hip.so")
ARCH = os.environ.get("TORCHWI_AMD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build torchwi_amd)")


def command(out=OUT):
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-fno-gpu-flush-denormals-to-zero",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I", os.path.join(ROOT, "include"),
        "-o", out, SRC,
    ]


def build(verbose=False):
    """Compile the library (always; callers decide whether to rebuild)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtorchwi_hip.so")
    os.replace(tmp, OUT)
    return OUT


def stale():
    if not os.path.exists(OUT):
        return True
    t = os.path.getmtime(OUT)
    return any(os.path.getmtime(p) > t for p in DEPS)


if __name__ == "__main__":
    print(build(verbose=True))
