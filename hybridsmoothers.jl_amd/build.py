"""Build recipe of libhsmooth.so (gfx950 only), in-tree.

    python hybridsmoothers.jl_amd/build.py            # incremental
    python hybridsmoothers.jl_amd/build.py --force    # rebuild everything

Every csrc/*.hip file is compiled by hipcc into an object under _build/ and the
objects are linked into hybridsmoothers/libhsmooth.so, the file the Python
package (and a Julia ccall wrapper) loads.  Flags:
  --offload-arch=gfx950          the only target
  -ffp-contract=off              no implicit FMA: sums round like the reference's
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                 no packed FP32 instructions
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "hybridsmoothers", "libhsmooth.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-pthread",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I", INCLUDE,
]


def _deps(src):
    hdrs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(INCLUDE, "*.h"))
    return [src] + hdrs


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, force):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if not force and not _stale(obj, _deps(src)):
        return obj, None
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return obj, None


CHECK_SRC = os.path.join(CSRC, "check", "hs_check.hip")
CHECK_OUT = os.path.join(HERE, "hybridsmoothers", "libhs_check.so")


def build_check(force: bool = False) -> str:
    """libhs_check.so: host builds of the sweeps' arithmetic helpers for the CPU
    tests (hs_check.hip), with the library's flags."""
    deps = _deps(CHECK_SRC)
    if force or _stale(CHECK_OUT, deps):
        tmp = CHECK_OUT + ".tmp"
        cmd = [HIPCC, *FLAGS, "-shared", CHECK_SRC, "-o", tmp]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        shutil.move(tmp, CHECK_OUT)
    return CHECK_OUT


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, force), srcs))
    errs = [e for _, e in results if e]
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))
    objs = [o for o, _ in results]
    if force or _stale(OUT, objs):
        tmp = OUT + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        shutil.move(tmp, OUT)
    build_check(force)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    print(build(a.force, a.j))
    sys.exit(0)
