"""Builds libfsph.so in-tree from flecsph_amd/csrc/*.hip for gfx950 with
hipcc (objects under flecsph_amd/_build/, compiled in parallel)."""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfsph.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FSPH_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions off (v_pk_* FP32 forms are avoided)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", SRC,
    "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(SRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o",
           tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, OUT)
    build_driver()
    if verbose:
        print(f"built {OUT} and {DRIVER}")
    return OUT


DRIVER = os.path.join(HERE, "fsph_hydro_driver")
FUNCTOR_CHECK = os.path.join(HERE, "fsph_functor_check")


def build_driver():
    """The C++ programs over host/body_system.hpp (the hydro driver loop
    and the host-functor example), linked to libfsph.so (host compiler
    only; the device code is in the library)."""
    for src, out in (("hydro_driver.cpp", DRIVER),
                     ("functor_check.cpp", FUNCTOR_CHECK)):
        cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
               os.path.join(ROOT, "include"), "-I",
               os.path.join(HERE, "host"), os.path.join(HERE, "host", src),
               "-L", HERE, "-lfsph", "-Wl,-rpath,$ORIGIN", "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{src} build failed:\n{r.stderr[-4000:]}")
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
