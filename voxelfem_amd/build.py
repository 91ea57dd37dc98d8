"""Build the native parts of voxelfem_amd in-tree.

* ``libvoxelfem.so``   -- hand-written HIP kernels for gfx950 + host orchestration, exporting the C ABI
                          of ``include/voxelfem.h`` (hipcc).
* ``pyVoxelFEM*.so``   -- pybind11 module mirroring the reference's ``pyVoxelFEM`` API on top of the
                          C ABI (g++; no HIP headers, no torch types).
* ``pyOptimizer*.so``  -- pybind11 module with the reference's ``pyOptimizer.MMA`` (host C++, OpenMP).

Run ``python -m voxelfem_amd.build`` (or ``__graft_entry__.build()``). Every call recompiles.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["vf_kernels_ops.hip", "vf_kernels_tiled.hip", "vf_kernels_walsh.hip", "vf_kernels_setup.hip",
               "vf_kernels_chol.hip", "vf_kernels_vec.hip", "vf_kernels_l1mfma.hip", "vf_planar.hip", "vf_kernels_filters.hip", "vf_host.hip"]
ARCH = os.environ.get("VF_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are disabled for every HIP translation unit
# code object v5 loads under both the system HIP runtime and the one bundled with torch
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-mcode-object-version=5", "-Xclang",
             "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
             "-Wno-unused-variable"]

LIB_NAME = "libvoxelfem.so"


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops' is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = os.path.join(BUILD, LIB_NAME + ".tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-mcode-object-version=5", "-shared", "-fPIC", *objs, "-o", tmp])
    shutil.move(tmp, os.path.join(HERE, LIB_NAME))
    return os.path.join(HERE, LIB_NAME)


def build_pybind(verbose=False):
    import pybind11

    src = os.path.join(CSRC, "pyvoxelfem.cpp")
    out = os.path.join(HERE, "pyVoxelFEM" + ext_suffix())
    tmp = os.path.join(BUILD, "pyVoxelFEM.so.tmp")
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-fvisibility=hidden", "-I", py_inc, "-I",
           pybind11.get_include(), "-I", INCLUDE, src, "-L", HERE, "-lvoxelfem", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    o = _run(cmd)
    if verbose and o.strip():
        print(o)
    shutil.move(tmp, out)
    return out


def build_optimizer(verbose=False):
    """pyOptimizer: the host-side MMA optimiser (no GPU code)."""
    import pybind11

    src = os.path.join(CSRC, "pyoptimizer.cpp")
    out = os.path.join(HERE, "pyOptimizer" + ext_suffix())
    os.makedirs(BUILD, exist_ok=True)
    tmp = os.path.join(BUILD, "pyOptimizer.so.tmp")
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-O3", "-shared", "-fPIC", "-std=c++17", "-fopenmp", "-fvisibility=hidden", "-I", py_inc, "-I",
           pybind11.get_include(), src, "-o", tmp]
    o = _run(cmd)
    if verbose and o.strip():
        print(o)
    shutil.move(tmp, out)
    return out


def build(verbose=False):
    lib = build_library(verbose=verbose)
    mod = build_pybind(verbose=verbose)
    opt = build_optimizer(verbose=verbose)
    return lib, mod, opt


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print("built", p)
