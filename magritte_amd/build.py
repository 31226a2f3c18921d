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


def core_path():
    return os.path.join(HERE, "core" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_host(verbose=False):
    """g++ build of the host layer into the extension module `core`."""
    import pybind11
    inc = ["-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"]]

    def comp(unit):
        obj = os.path.join(BUILD, "host_" + unit.replace(".cpp", ".o"))
        cmd = ["g++", *HOST_FLAGS, *inc, "-c", os.path.join(HOST, unit), "-o", obj]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"compile of {unit} failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        return obj, res.stderr

    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(HOST_UNITS)) as ex:
        results = list(ex.map(comp, HOST_UNITS))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    out = core_path()
    tmp = out + ".tmp"
    cmd = ["g++", "-shared", "-o", tmp, *[r[0] for r in results], "-L" + HERE, "-lmagritte_hip",
           "-Wl,-rpath,$ORIGIN"]
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
