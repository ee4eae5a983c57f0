"""Build librbhip.so in-tree with hipcc for gfx950.

    python -m redbirdpy_amd.build
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["rb_mesh.hip", "rb_assemble.hip", "rb_krylov.hip", "rb_jac.hip", "rb_gram.hip", "rb_widefield.hip", "rb_bicgstab.hip", "rb_chol.hip"]
OUT = os.path.join(HERE, "librbhip.so")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def build_lib(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building librbhip.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_lib(verbose=True))
