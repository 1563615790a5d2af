"""Build the MI355X FX engine in-tree.

Outputs (git-ignored, shipped with the tree):
  difx_tests_amd/lib/libdifxfx.so   HIP kernels + native C API + kernel.h ABI
  difx_tests_amd/lib/fxmain         drop-in host program (reference main.c)
  difx_tests_amd/lib/fxbatch        uneven-subband host program (native API)
  oracle/_build/libfxoracle.so      CPU oracle (test infrastructure only)

Usage: python -m difx_tests_amd.build [--no-oracle]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "difx_tests_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["fx_fstage.hip", "fx_xengine.hip", "fx_single.hip", "fx_api.cpp", "fx_kernelh.cpp", "fx_noalloc.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def build_library(jobs: int = 8) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"]
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(cmds)))) as ex:
        list(ex.map(_run, cmds))
    so = os.path.join(LIBDIR, "libdifxfx.so")
    tmp = so + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, so)
    return so


def build_host() -> str:
    exe = os.path.join(LIBDIR, "fxmain")
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", f"-I{INCLUDE}",
          os.path.join(PKG, "host", "fxmain.c"), "-o", exe,
          f"-L{LIBDIR}", "-ldifxfx", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-lm"])
    return exe


def build_batch_host() -> str:
    exe = os.path.join(LIBDIR, "fxbatch")
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", f"-I{INCLUDE}",
          os.path.join(PKG, "host", "fxbatch.c"), "-o", exe,
          f"-L{LIBDIR}", "-ldifxfx", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-lm"])
    return exe


def build_oracle() -> str:
    outdir = os.path.join(ORACLE, "_build")
    os.makedirs(outdir, exist_ok=True)
    so = os.path.join(outdir, "libfxoracle.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          os.path.join(ORACLE, "fx_oracle.c"), "-o", so, "-lm"])
    return so


def build_test_drivers() -> str:
    """C driver of the per-(antenna, FFT) kernel.h functions (tests only)."""
    outdir = os.path.join(ROOT, "tests", "_build")
    os.makedirs(outdir, exist_ok=True)
    exe = os.path.join(outdir, "kernelh_driver")
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", f"-I{INCLUDE}", os.path.join(ROOT, "tests", "kernelh_driver.c"),
          "-o", exe, f"-L{LIBDIR}", "-ldifxfx", "-Wl,-rpath,$ORIGIN/../../difx_tests_amd/lib",
          "-Wl,-rpath,/opt/rocm/lib", "-lm"])
    return exe


def build_reference_host() -> None:
    """The reference's main.c linked against this library, when the reference is present."""
    if os.path.exists("/root/reference/main.c"):
        _run(["bash", os.path.join(ORACLE, "build_ref_host.sh")])


def build_all(oracle: bool = True) -> None:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_library()
    build_host()
    build_batch_host()
    if oracle:
        build_oracle()
        build_test_drivers()
        build_reference_host()


if __name__ == "__main__":
    build_all(oracle="--no-oracle" not in sys.argv)
    print("built", os.path.join(LIBDIR, "libdifxfx.so"))
