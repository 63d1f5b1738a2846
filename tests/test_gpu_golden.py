"""The HIP path (through the C ABI) reproduces every committed golden
fixture bit-exactly: migrate, exchange_halo with registered fields,
reverse_halo and reverse_exchange (Sum, Max)."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import CASES
from tests.gpu_util import F64, I32, assert_bits_equal, host, say

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_matches_fixture(gpu, name):
    torch = gpu
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig, DeviceParticles, LoopbackHalo
    say("golden", name)
    fx = np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False)
    n, L, lower = int(fx["n"]), float(fx["L"]), float(fx["lower"])
    g = int(fx["ghosts"])
    box = BoxConfig(lower=(lower,) * 3, upper=(lower + L,) * 3,
                    periodic=tuple(bool(v) for v in fx["periodic"]), width=float(fx["width"]),
                    edges_and_corners=bool(fx["edges"]))
    cap = n + g + 32
    p = DeviceParticles(cap)
    cuda = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    for k in F64 + I32:
        getattr(p, k)[:n] = cuda(fx[f"in_{k}"])
    for a, k in enumerate(("x", "y", "z")):
        getattr(p, k)[:n] += cuda(fx["in_disp"][a])
    halo = LoopbackHalo(box)
    halo.migrate(p, n)
    for k in ("x", "y", "z"):
        assert_bits_equal(host(getattr(p, k)[:n]), fx[f"migrated_{k}"], f"migrated {k}")
    fields = []
    with_fields = "in_density" in fx.files
    if with_fields:
        dens = torch.zeros(cap, dtype=torch.float64, device="cuda")
        vec = torch.zeros((cap, 3), dtype=torch.float64, device="cuda")
        dens[:n] = cuda(fx["in_density"])
        vec[:n] = cuda(fx["in_vec"])
        fields = [dens, vec]
    assert halo.exchange_halo(p, n, fields=fields) == g
    for k in F64 + I32:
        assert_bits_equal(host(getattr(p, k)[:n + g]), fx[f"halo_{k}"], f"halo {k}")
    assert_bits_equal(host(halo.ghost_owners(g)), fx["ghost_owners"], "ghost owners")
    if with_fields:
        assert_bits_equal(host(dens[:n + g]), fx["halo_density"], "halo density")
        assert_bits_equal(host(vec[:n + g]).reshape(-1), fx["halo_vec"].reshape(-1), "halo vec")
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(p, k)[:n + g] = cuda(fx["in_forces"][a])
    halo.reverse_halo(p, n)
    for k in ("fx", "fy", "fz"):
        assert_bits_equal(host(getattr(p, k)[:n]), fx[f"reversed_{k}"], f"reversed {k}")
    if with_fields:
        dens[:n + g] = cuda(fx["in_contrib"])
        halo.reverse_exchange([(dens, _abi.SUM)], n)
        assert_bits_equal(host(dens[:n]), fx["reversed_density"], "reversed density")
        vec[:n + g] = cuda(fx["in_vmax"])
        halo.reverse_exchange([(vec, _abi.MAX)], n)
        assert_bits_equal(host(vec[:n]).reshape(-1), fx["reversed_vec_max"].reshape(-1),
                          "reversed vec max")
