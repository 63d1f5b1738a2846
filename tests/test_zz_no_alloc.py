"""dsdx_halo_build_noalloc (plan state in the library's fixed device array)
against dsdx_halo_build (plan buffers) on the same inputs, through the C ABI:
every particle array, every registered field, the ghost count and the ghost
owners bit-exact, on two calls in a row, the second with a smaller box; and
both against the oracle.  Also: a call whose ghosts do not fit the storage
writes none and reports DSDX_ECAPACITY with the count, and the calls after it
are still exact."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import assert_bits_equal, compare_particles, host, say, synth, to_device, to_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dsdecompx_amd import _abi
    _abi.lib()
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


def _noalloc(box, dev, n, capacity, fields, owners):
    import torch
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import _field_array
    arr, nf = _field_array(fields)
    g = C.c_uint64(0)
    pc = dev.c()
    rc = _abi.lib().dsdx_halo_build_noalloc(C.byref(box.c()), C.byref(pc), arr, nf, n, capacity,
                                            owners.data_ptr() if owners is not None else None,
                                            C.c_void_p(torch.cuda.current_stream().cuda_stream),
                                            C.byref(g))
    return rc, int(g.value)


def _case(gpu, n, L, seed, w=2.8):
    import torch
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    say("noalloc build", n, L, seed)
    d = synth(n, L, seed)
    cap = 3 * n + 64
    vec = np.random.default_rng(seed + 1).uniform(-1, 1, (n, 3))
    # the existing entry point: a plan with its own buffers
    ref = to_device(d, cap)
    ref_vec = torch.zeros((cap, 3), dtype=torch.float64, device="cuda")
    ref_vec[:n] = torch.from_numpy(vec).cuda()
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    g_ref = halo.exchange_halo(ref, n, fields=[ref_vec])
    ref_owners = host(halo.ghost_owners(g_ref))
    # the new one, same inputs, fresh storage
    new = to_device(d, cap)
    new_vec = torch.zeros((cap, 3), dtype=torch.float64, device="cuda")
    new_vec[:n] = torch.from_numpy(vec).cuda()
    owners = torch.full((cap - n,), -7, dtype=torch.int32, device="cuda")
    rc, g = _noalloc(BoxConfig(upper=(L,) * 3, width=w), new, n, cap, [new_vec], owners)
    _abi.check(rc, "noalloc build")
    assert g == g_ref
    for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "charge", "type",
              "global_id"):
        assert_bits_equal(host(getattr(new, k)[:n + g]), host(getattr(ref, k)[:n + g]), k)
    assert_bits_equal(host(new_vec[:n + g]), host(ref_vec[:n + g]), "field")
    assert_bits_equal(host(owners[:g]), ref_owners, "ghost owners")
    # and the oracle
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
    op = to_oracle(d, cap)
    assert ora.exchange_halo(op, n) == g
    compare_particles(new, op, n + g, "noalloc vs oracle")
    assert_bits_equal(host(owners[:g]), ora.ghost_owners(), "ghost owners vs oracle")
    halo.close()
    return g


def test_noalloc_equals_plan_build_twice(gpu):
    g1 = _case(gpu, 32768, 32.0, 3)
    g2 = _case(gpu, 5000, 18.0, 4)   # smaller: the array keeps the first call's contents
    assert g1 > g2 > 0


def test_noalloc_overflow_then_exact(gpu):
    import torch
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig
    say("noalloc overflow")
    n, L = 20000, 27.0
    d = synth(n, L, 8)
    small = n + 100
    dev = to_device(d, small)
    for k in ("x", "vx", "global_id"):
        getattr(dev, k)[n:] = -3
    rc, g = _noalloc(BoxConfig(upper=(L,) * 3, width=2.8), dev, n, small, [], None)
    assert rc == _abi.ECAPACITY and g > 100
    assert b"capacity" in _abi.lib().dsdx_last_error()
    for k in ("x", "vx", "global_id"):
        assert (host(getattr(dev, k)[n:]) == -3).all()   # no ghost written
    torch.cuda.synchronize()
    _case(gpu, 12000, 24.0, 9)
