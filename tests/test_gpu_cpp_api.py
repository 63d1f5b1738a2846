"""Runs the C++ API test program (tests/cpp/decomp_api_test.cpp): Decomp /
ParticleInfo on device memory checked like the reference's example and
sanity programs and bit-exactly against the oracle, step by step."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_decomp_api_program():
    exe = os.path.join(HERE, "cpp", "bin", "decomp_api_test")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run __graft_entry__.build() first")
    print("[case] decomp_api_test", flush=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "all checks passed" in r.stdout
