"""Writes tests/golden/ref_grid_2x2x2.npz from the reference itself (TEST
INFRASTRUCTURE, run in the build container only, where /root/reference and
MPICH exist): builds the reference's Host library and this project's fixture
writer tests/golden/ref_grid_fixture.cpp with oracle/ref.mk, runs it on 8 MPI
ranks (a 2x2x2 periodic process grid, the reference unmodified) and stores
what every rank holds after each operation as arrays named
"r<rank>_s<step>_<name>" (start arrays as step 0), plus the box.  The GPU runs
and the tests read only the .npz.

    python tests/golden/make_ref_fixtures.py            (from the repository root)
    python tests/golden/make_ref_fixtures.py --time     (also times the reference)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
MPIEXEC = "/opt/conda/bin/mpiexec"
ENV = dict(os.environ, LD_LIBRARY_PATH="/usr/lib/x86_64-linux-gnu:/opt/conda/lib")
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_grid_fixture")
OUT = os.path.join(HERE, "ref_grid_2x2x2.npz")
INTS = {"global_id", "type"}


def build():
    if not os.path.isdir(REF):
        raise SystemExit(f"{REF} is not here: the fixture can only be made where the reference is")
    subprocess.run(["make", "-s", "-j8", "-f", os.path.join(ROOT, "oracle", "ref.mk")], check=True)


def parse(path, rank):
    arrays = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            step, name, count = int(parts[0]), parts[1], int(parts[2])
            values = parts[3:3 + count]
            if name.split("_", 1)[1] in INTS:
                arr = np.array([int(v) for v in values], np.int32)
            else:
                arr = np.array([float.fromhex(v) for v in values], np.float64)
            arrays[f"r{rank}_s{step}_{name}"] = arr
    return arrays


def make():
    build()
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.run([MPIEXEC, "-n", "8", EXE, "fixture", tmp], check=True, env=ENV, cwd="/tmp",
                       timeout=300)
        arrays = {}
        for r in range(8):
            arrays.update(parse(os.path.join(tmp, f"rank{r}.txt"), r))
    arrays["box_lower"] = np.array([-3.7, 1.3, 0.25])
    arrays["box_upper"] = np.array([12.1, 20.9, 9.85])
    arrays["box_width"] = np.array([1.1])
    arrays["dims"] = np.array([2, 2, 2], np.int32)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {len(arrays)} arrays")


def time_reference():
    """The reference Host path at 8 ranks on the uniform workloads (the
    communication_test step), slowest rank's mean ms per step."""
    build()
    import json
    out = []
    for n, L, steps in ((32768, 32.0, 50), (16777216, 256.0, 5)):
        r = subprocess.run([MPIEXEC, "-n", "8", EXE, "time", str(n), repr(L), "2.8", "0.1", str(steps)],
                           check=True, env=ENV, cwd="/tmp", stdout=subprocess.PIPE, text=True,
                           timeout=1800)
        out.append(json.loads(r.stdout.strip().splitlines()[-1]))
        print(out[-1], flush=True)
    return out


if __name__ == "__main__":
    make()
    if "--time" in sys.argv:
        time_reference()
