"""Build libgpsat.so (HIP for gfx950 + host C++) and the gpupsat CLI in-tree with hipcc.

Objects go to gpupsat_amd/_build/ (git-ignored); the shared library and the CLI land
next to this file so they travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpsat.so")
CLI = os.path.join(HERE, "bin", "gpupsat")
ARCH = os.environ.get("GPSAT_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed fp32 forms are kept out of every device build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer (device code is built as usual):
# each -fsanitize= goes to the host pass only.
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
SAN_OBJ = os.path.join(HERE, "_build_san")
SAN_LIB = os.path.join(SAN_OBJ, "libgpsat.so")


def _compile(src: str, sanitize: bool = False) -> str:
    obj = os.path.join(SAN_OBJ if sanitize else OBJ, os.path.basename(src) + ".o")
    cmd = [_hipcc(), *COMMON, *(SAN_HOST if sanitize else []), "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cli_src = os.path.join(HERE, "cli", "gpupsat.cpp")
    if os.path.exists(cli_src):
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        cmd = [_hipcc(), "-O2", "-std=c++17", "-x", "hip", f"--offload-arch={ARCH}",
               "-I" + os.path.join(ROOT, "include"), cli_src, "-o", CLI + ".tmp",
               "-L" + HERE, "-lgpsat", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(CLI + ".tmp", CLI)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


def build_sanitized(jobs: int | None = None) -> str:
    """libgpsat.so with sanitized host code, in gpupsat_amd/_build_san/ (CPU testing only)."""
    os.makedirs(SAN_OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, True), srcs))
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-shared-libasan",
           "-Xarch_host", "-fsanitize=address,undefined", "-o", SAN_LIB, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return SAN_LIB


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized())
    else:
        build(verbose=True)
