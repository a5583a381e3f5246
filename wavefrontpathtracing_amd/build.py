"""Build the HIP library (libwf.so) in-tree for gfx950.

Each source is compiled to an object under ``_build/`` in parallel, then linked
with hipcc.  Device code is compiled without floating-point contraction,
with correctly rounded FP32 divide/sqrt and denormals preserved, so that the
closest-hit arithmetic matches an IEEE-754 FP32 evaluation of the same
expressions on the host.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libwf.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fvisibility=hidden", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
CXX_FLAGS = []


def _sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-c", src, "-o", obj] + COMMON + HIP_FLAGS
    else:
        cmd = ["g++", "-c", src, "-o", obj] + COMMON + CXX_FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, len(srcs))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)


SAN_FLAGS = ["-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=undefined",
             "-fno-omit-frame-pointer", "-g"]


def build_sanitized(out_dir: str, verbose: bool = False) -> str:
    """A copy of the library whose host C++ sources (the BLAS/TLAS builders,
    camera, scene generators, image files) are compiled with AddressSanitizer
    and UndefinedBehaviorSanitizer, linked with the regular HIP objects.  Load
    it with WF_LIB=<path> and LD_PRELOAD of gcc's libasan (the CPU tests do)."""
    os.makedirs(out_dir, exist_ok=True)
    hip_objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in _sources() if s.endswith(".hip")]
    if not all(os.path.exists(o) for o in hip_objs):
        build()
    objs = []
    for src in _sources():
        if not src.endswith(".cpp"):
            continue
        obj = os.path.join(out_dir, os.path.basename(src) + ".o")
        cmd = ["g++", "-c", src, "-o", obj] + [f for f in COMMON if f != "-O3"] + ["-O1"] + SAN_FLAGS
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    lib = os.path.join(out_dir, "libwf.so")
    cmd = ["g++", "-shared", "-fPIC", "-o", lib] + objs + hip_objs + SAN_FLAGS + \
          ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {lib}")
    return lib
