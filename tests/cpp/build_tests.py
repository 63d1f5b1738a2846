"""Builds the C++ test programs of tests/cpp/ (run by tests/test_gpu_cpp_api.py)
against the in-tree library and the CPU oracle (linked as the checker)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
MPI_INCLUDE = os.environ.get("DSDX_MPI_INCLUDE", "/opt/conda/include")
PROGRAMS = ["decomp_api_test"]


def build() -> list:
    os.makedirs(BIN, exist_ok=True)
    out = []
    for name in PROGRAMS:
        exe = os.path.join(BIN, name)
        cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
               "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "include", "dsdecompx"),
               "-I", os.path.join(ROOT, "oracle"), "-isystem", MPI_INCLUDE,
               "-DMPICH_SKIP_MPICXX", "-DOMPI_SKIP_MPICXX",
               os.path.join(HERE, name + ".cpp"),
               "-L", os.path.join(ROOT, "dsdecompx_amd", "lib"), "-ldsdecompx_amd",
               "-L", os.path.join(ROOT, "oracle"), "-loracle_dsdx",
               "-Wl,-rpath,$ORIGIN/../../../dsdecompx_amd/lib",
               "-Wl,-rpath,$ORIGIN/../../../oracle", "-o", exe]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {name} failed:\n{r.stdout}")
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build())
