"""Build librbhip.so in-tree with hipcc for gfx950.

    python -m redbirdpy_amd.build

Each HIP source is compiled to its own object (in parallel; every
translation unit carries its own device code, no relocatable device code),
then the objects are linked into one shared library.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
SOURCES = ["rb_mesh.hip", "rb_meshprep.hip", "rb_assemble.hip", "rb_krylov.hip", "rb_jac.hip", "rb_gram.hip", "rb_widefield.hip",
           "rb_bicgstab.hip", "rb_chol.hip"]
OUT = os.path.join(HERE, "librbhip.so")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def _stale(src, obj):
    """An object is stale when it or its source is missing or older than the source or a header."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps.append(os.path.join(os.path.dirname(HERE), "include", "rbhip.h"))
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build_lib(verbose=False, jobs=None, force=True):
    """Compile every source (force) or only those whose object is stale, then link."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    srcs = [s for s, o in zip(SOURCES, objs) if force or _stale(os.path.join(CSRC, s), o)]
    cmds = [[hipcc] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", os.path.join(OBJ, os.path.splitext(s)[0] + ".o")]
            for s in srcs]
    jobs = jobs or min(max(len(cmds), 1), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        rcs = list(ex.map(lambda c: _run(c, verbose), cmds))
    if any(rcs):
        raise RuntimeError("hipcc failed compiling %s" % ", ".join(s for s, rc in zip(srcs, rcs) if rc))
    tmp = OUT + ".tmp"
    if _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose):
        raise RuntimeError("hipcc failed linking librbhip.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_lib(verbose=True, force="--changed" not in sys.argv))
