"""Build recipe of the gfx950 library libmagritte_hip.so (in-tree).

Usage: python -m magritte_amd.build   (or __graft_entry__.build()).

Also builds the host layer (csrc/host: Io, Model, Simulation over the C-ABI)
with g++ into the Python module magritte_amd.core, linked to the library.

Every HIP translation unit is compiled for gfx950 only, with packed FP32
instructions disabled.  The set-up and tracer units restate FP64 expressions
of the reference and are compiled without floating-point contraction so that
they round exactly as the reference's scalar code; the radiation-field unit
marks its exact FP64 expressions with `#pragma clang fp contract(off)` and is
otherwise free to fuse.
"""
import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmagritte_hip.so")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fno-gpu-rdc", "-munsafe-fp-atomics"]

UNITS = {
    "mgr_context.hip": ["-ffp-contract=off"],
    "mgr_trace.hip": ["-ffp-contract=off"],
    "mgr_radiation.hip": [],
    "mgr_levels.hip": [],
    "mgr_mlp.hip": [],
}


def hipcc():
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(unit, extra):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
    cmd = [hipcc(), *COMMON, *extra, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile of {unit} failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj, res.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        results = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    objs = [r[0] for r in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,-soname,libmagritte_hip.so",
           "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


HOST = os.path.join(CSRC, "host")
HOST_UNITS = ["io_text.cpp", "model.cpp", "image.cpp", "simulation.cpp", "core.cpp"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-fvisibility=hidden"]


def hdf5_prefix():
    """Install prefix of the HDF5 C library, or None: the IoHDF5 back end is
    built only where the library is."""
    for prefix in (os.environ.get("HDF5_DIR", ""), "/opt/conda", "/usr", "/usr/local"):
        if prefix and os.path.exists(os.path.join(prefix, "include", "hdf5.h")):
            for libdir in ("lib", "lib64", "lib/x86_64-linux-gnu"):
                if os.path.exists(os.path.join(prefix, libdir, "libhdf5.so")):
                    return prefix, os.path.join(prefix, libdir)
    return None


def core_path():
    return os.path.join(HERE, "core" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_host(verbose=False):
    """g++ build of the host layer into the extension module `core`."""
    import pybind11
    inc = ["-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"]]
    units, link_extra = list(HOST_UNITS), []
    hdf5 = hdf5_prefix()
    if hdf5:
        prefix, libdir = hdf5
        units.append("io_hdf5.cpp")
        inc += ["-DMGR_WITH_HDF5"]
        link_extra = ["-L" + libdir, "-lhdf5", "-Wl,-rpath," + libdir]

    def comp(unit):
        obj = os.path.join(BUILD, "host_" + unit.replace(".cpp", ".o"))
        extra = ["-isystem", os.path.join(hdf5[0], "include")] if (hdf5 and unit == "io_hdf5.cpp") else []
        cmd = ["g++", *HOST_FLAGS, *inc, *extra, "-c", os.path.join(HOST, unit), "-o", obj]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"compile of {unit} failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        return obj, res.stderr

    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(units)) as ex:
        results = list(ex.map(comp, units))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    out = core_path()
    tmp = out + ".tmp"
    cmd = ["g++", "-shared", "-o", tmp, *[r[0] for r in results], "-L" + HERE, "-lmagritte_hip",
           "-Wl,-rpath,$ORIGIN", *link_extra]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link of core failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    lib = build(verbose)
    return lib, build_host(verbose)


if __name__ == "__main__":
    print(build_all(verbose=True))
