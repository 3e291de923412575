"""Build libspt_hip.so in-tree with hipcc for gfx950.

Flags: -ffp-contract=off keeps every unfused expression rounding like the
NumPy reference (hot loops fuse explicitly with fma()); packed FP32
instructions are disabled (-packed-fp32-ops)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "spt.hip")
OUT = os.path.join(HERE, "libspt_hip.so")
ARCH = os.environ.get("SPT_OFFLOAD_ARCH", "gfx950")


def hipcc_cmd(out=OUT, extra=()):
    return [
        "hipcc", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        SRC, "-o", out, "-lrocsolver", "-lrocblas", *extra,
    ]


def build(verbose=False, extra=()):
    tmp = OUT + ".tmp"
    cmd = hipcc_cmd(tmp, extra)
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-8000:]}")
    if verbose and r.stderr:
        print(r.stderr[-20000:], file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True, extra=tuple(sys.argv[1:]))
