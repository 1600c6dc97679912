"""Build the HIP library (and the C++ class surface) in-tree.

Outputs go to visualtur_amd/_build/:
  libvt_hip.so      C ABI of include/visualtur_amd.h (kernels + host runtime)
  libvisualtur.so   the reference-compatible C++ classes on top of the C ABI
  deps/             symlinks to the image's libhdf5 (RUNPATH $ORIGIN/deps),
                    so /opt/conda/lib never lands in a library search path.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
INCLUDE = ROOT / "include"
HDF5_INC = Path("/opt/conda/include")
HDF5_LIB = Path("/opt/conda/lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
INCS = [f"-I{INCLUDE}", f"-I{CSRC}", f"-I{HDF5_INC}"]

HIP_SOURCES = ["vt_kernels.hip", "vt_build.hip", "vt_mlp.hip", "vt_runtime.cpp", "vt_prefetch.cpp",
               "vt_serve.cpp", "vt_frames.cpp"]
HOST_SOURCES = ["vt_hdf5.cpp"]
CLASS_SOURCES = ["vt_classes.cpp"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:4])} ...")


def _deps() -> Path:
    deps = BUILD / "deps"
    deps.mkdir(parents=True, exist_ok=True)
    for name in ("libhdf5.so.103",):
        link = deps / name
        if not link.exists():
            link.symlink_to(HDF5_LIB / name)
    dev = deps / "libhdf5.so"
    if not dev.exists():
        dev.symlink_to(HDF5_LIB / "libhdf5.so.103")
    return deps


def build(verbose: bool = False) -> Path:
    """Compile every source for gfx950; returns the path of libvt_hip.so."""
    BUILD.mkdir(parents=True, exist_ok=True)
    obj = BUILD / "obj"
    obj.mkdir(exist_ok=True)
    deps = _deps()
    jobs = []
    for src in HIP_SOURCES:
        jobs.append([HIPCC, *COMMON, *HIP, "-x", "hip", *INCS, "-c", str(CSRC / src),
                     "-o", str(obj / (src + ".o"))])
    classes = [s for s in CLASS_SOURCES if (CSRC / s).exists()]
    for src in HOST_SOURCES + classes:
        jobs.append([HIPCC, *COMMON, *INCS, "-c", str(CSRC / src), "-o", str(obj / (src + ".o"))])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    lib = BUILD / "libvt_hip.so"
    _run([HIPCC, *HIP, "-shared", "-fPIC", "-o", str(lib),
          *[str(obj / (s + ".o")) for s in HIP_SOURCES + HOST_SOURCES],
          f"-L{deps}", "-lhdf5", "-Wl,-rpath,$ORIGIN/deps", "-Wl,-z,defs"], verbose)
    if not classes:
        return lib
    cls = BUILD / "libvisualtur.so"
    _run([HIPCC, "-shared", "-fPIC", "-o", str(cls), *[str(obj / (s + ".o")) for s in classes],
          f"-L{BUILD}", "-lvt_hip", "-Wl,-rpath,$ORIGIN", f"-L{deps}", "-lhdf5",
          "-Wl,-rpath,$ORIGIN/deps"], verbose)
    # the headless viewer (drawPixel.cu without a window) on the class surface
    _run([HIPCC, *COMMON, *INCS, str(CSRC / "vt_view.cpp"), "-o", str(BUILD / "vt_view"), f"-L{BUILD}", "-lvisualtur",
          "-lvt_hip", "-Wl,-rpath,$ORIGIN", f"-L{deps}", "-lhdf5", "-Wl,-rpath,$ORIGIN/deps"], verbose)
    # C++ drivers of the class surface used by the tests
    for drv in sorted((ROOT / "tests" / "cpp").glob("*.cpp")):
        _run([HIPCC, *COMMON, *INCS, str(drv), "-o", str(BUILD / drv.stem), f"-L{BUILD}", "-lvisualtur",
              "-lvt_hip", "-Wl,-rpath,$ORIGIN", f"-L{deps}", "-lhdf5", "-Wl,-rpath,$ORIGIN/deps"], verbose)
    return lib


def library_path() -> Path:
    return BUILD / "libvt_hip.so"


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
