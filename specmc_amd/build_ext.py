"""Build the in-tree HIP library ``specmc_amd/_lib/libspecmc_hip.so`` for gfx950.

    python -m specmc_amd.build_ext

The library is compiled with hipcc directly (no torch extension machinery):
it exports only the C ABI of include/specmc_hip.h and is loaded with ctypes.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "csrc", f)
       for f in ("specmc_kernels.hip", "dense_lu.hip", "schur_apply.hip", "gj_mfma.hip",
                          "general_state.hip")]
OUT_DIR = os.path.join(HERE, "_lib")
OUT = os.path.join(OUT_DIR, "libspecmc_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-I", os.path.join(ROOT, "include"),
    # packed FP32 instructions are disabled for every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast",
    # the register-resident LU sub-panel (k_lud_panel, 4 rows x 16 columns per
    # thread) is a fully unrolled column loop; at the default limit it stays
    # rolled and its array goes to scratch memory
    "-mllvm", "-pragma-unroll-threshold=200000",
]


def build(verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + SRC + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building %s" % OUT)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
