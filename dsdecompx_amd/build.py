"""Builds the gfx950 shared library of the package in-tree with hipcc.

``python -m dsdecompx_amd.build`` (or ``build()``) compiles every HIP/C++
source under ``csrc/`` into ``dsdecompx_amd/lib/libdsdecompx_amd.so``.
FP contraction is off so that every FP64 operation rounds on its own, as in
the reference Host build.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libdsdecompx_amd.so")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# The public C++ API includes <mpi.h> for the MPI_Comm type of the Decomp
# constructor; the library calls no MPI function, so only the header is used.
MPI_INCLUDE = os.environ.get("DSDX_MPI_INCLUDE", "/opt/conda/include")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-I", INCLUDE, "-I", os.path.join(INCLUDE, "dsdecompx"), "-I", CSRC,
          "-isystem", MPI_INCLUDE, "-DMPICH_SKIP_MPICXX", "-DOMPI_SKIP_MPICXX"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    out = []
    for root, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".hip", ".cpp")):
                out.append(os.path.join(root, f))
    return sorted(out)


def _compile(src: str, obj: str, jobs_log: list) -> subprocess.Popen:
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", src, "-o", obj]
    jobs_log.append(" ".join(cmd))
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    objdir = os.path.join(ROOT, "_build", "obj")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    log: list = []
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.relpath(src, CSRC).replace(os.sep, "_") + ".o")
        objs.append(obj)
        procs.append((src, _compile(src, obj, log)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if verbose and out:
            sys.stdout.write(out)
        if p.returncode != 0:
            failed.append((src, out))
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


APPS = os.path.join(PKG, "apps")
BINDIR = os.path.join(PKG, "bin")


def build_apps() -> list:
    """Programs of apps/ (the C++ benchmark driver), linked to the library."""
    os.makedirs(BINDIR, exist_ok=True)
    out = []
    for f in sorted(os.listdir(APPS)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(BINDIR, f[:-4])
        cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", os.path.join(APPS, f), "-L", LIBDIR,
                                           "-ldsdecompx_amd", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {f} failed:\n{r.stdout}")
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_apps())
