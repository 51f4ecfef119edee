"""Builds the native libraries in-tree.

  ligmg_amd/libligmg_amd.so     HIP solver for gfx950 (hipcc)
  ligmg_amd/libligmg_graphs.so  host graph generators (g++, OpenMP)

Run as ``python -m ligmg_amd.build`` or through __graft_entry__.build().
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["prims.hip", "solve.hip", "setup.hip", "batch.hip", "solver.hip", "order.hip", "graphprep.hip", "tail.hip", "pool.hip", "pc.hip", "dense.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: sums and products round exactly as written, which the
# bit-exact setup relies on.  Packed FP32 instructions are disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
             "-Wno-unused-result", "-Wno-unused-variable"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_solver(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []

    def one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        out = _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    target = os.path.join(HERE, "libligmg_amd.so")
    tmp = target + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, target)
    return target


def build_graphs():
    target = os.path.join(HERE, "libligmg_graphs.so")
    tmp = target + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-shared",
          "-o", tmp, os.path.join(CSRC, "graphs.cpp")])
    os.replace(tmp, target)
    return target


def build_petsc_pc():
    """The PETSc "ligmg" PC (csrc/petsc/ligmg_pc.cpp), only where PETSc is
    installed: PETSC_DIR (and PETSC_ARCH) must point at it."""
    pdir = os.environ.get("PETSC_DIR")
    if not pdir:
        return None
    arch = os.environ.get("PETSC_ARCH", "")
    inc = [os.path.join(pdir, "include"), os.path.join(pdir, arch, "include")]
    lib = os.path.join(pdir, arch, "lib")
    target = os.path.join(HERE, "libligmg_petsc.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *[f"-I{i}" for i in inc],
          os.path.join(CSRC, "petsc", "ligmg_pc.cpp"), "-o", target, f"-L{lib}", "-lpetsc",
          f"-L{HERE}", "-lligmg_amd", f"-Wl,-rpath,{HERE}", f"-Wl,-rpath,{lib}"])
    return target


def build_all(verbose=False):
    g = build_graphs()
    s = build_solver(verbose)
    out = [g, s]
    p = build_petsc_pc()
    if p:
        out.append(p)
    return out


if __name__ == "__main__":
    for t in build_all(verbose="-v" in sys.argv):
        print("built", t)
