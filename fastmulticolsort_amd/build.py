"""Builds libfmcs.so in-tree with hipcc for gfx950.

Every source compiles to its own object in parallel (each .hip file is a
separate translation unit with its own code object), then one link.
incremental=True recompiles only objects older than their source or any
header; build() always recompiles everything.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libfmcs.so")

HIP_SOURCES = ["fmcs_kernels.hip", "fmcs_dd.hip", "fmcs_merge.hip", "fmcs_capi.hip", "fmcs_mlp.hip",
               "fmcs_hash.hip", "fmcs_noalloc.hip", "fmcs_bins.hip"]
HOST_SOURCES = ["fmcs_plan.cpp", "fmcs_composer.cpp", "fmcs_io.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _clean(stderr):
    return "\n".join(l for l in stderr.splitlines() if "packed-fp32-ops" not in l)


def _compile(src, obj):
    cmd = ["hipcc"] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, _clean(r.stderr)


def build_library(verbose=False, incremental=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    headers = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    hdr_time = max((os.path.getmtime(h) for h in headers), default=0.0)
    objs, todo = [], []
    for s in srcs:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs.append(o)
        if (not incremental or not os.path.exists(o)
                or os.path.getmtime(o) < max(os.path.getmtime(s), hdr_time)):
            todo.append((s, o))
    jobs = jobs or min(len(todo) or 1, 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, rc, err in ex.map(lambda so: _compile(*so), todo):
            if verbose:
                print("hipcc -c %s" % os.path.basename(src), flush=True)
            if rc != 0:
                raise RuntimeError(f"hipcc failed on {src} ({rc}):\n{err}")
            if verbose and err.strip():
                print(err, file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed ({r.returncode}):\n{_clean(r.stderr)}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True, incremental="--incremental" in sys.argv))
