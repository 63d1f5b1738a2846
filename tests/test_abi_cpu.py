"""CPU-side checks of the C ABI: the library loads and exports every function
that include/dsdx_hip.h declares, and argument validation errors come back as
status codes with messages (no GPU work is launched)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dsdx_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from dsdecompx_amd import _abi
    from dsdecompx_amd import build as b
    if not os.path.exists(b.LIB):
        b.build()
    return _abi.lib()


def test_every_declared_symbol_is_exported(lib):
    from dsdecompx_amd import _abi
    names = _declared("dsdx_hip.h")
    assert names, "no declarations found"
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_abi.exported_symbols()) == names


def test_validation_errors_without_device(lib):
    from dsdecompx_amd import _abi
    plan = C.c_void_p()
    assert lib.dsdx_plan_create(C.byref(plan)) == 0
    box = _abi.Box()
    box.upper[:] = [1.0, 1.0, 1.0]
    box.width = -1.0
    rc = lib.dsdx_migrate_wrap(C.byref(box), None, None, None, 0, C.c_void_p(1))
    assert rc == _abi.EINVAL and b"negative" in lib.dsdx_last_error()
    box.width = 0.5
    box.upper[0] = 0.0
    rc = lib.dsdx_migrate_wrap(C.byref(box), None, None, None, 0, C.c_void_p(1))
    assert rc == _abi.EINVAL and b"extent" in lib.dsdx_last_error()
    box.upper[0] = 1.0
    rc = lib.dsdx_migrate_wrap(C.byref(box), None, None, None, 0, None)
    assert rc == _abi.EINVAL and b"stream" in lib.dsdx_last_error()
    p = _abi.Particles()
    rc = lib.dsdx_halo_build(plan, C.byref(box), C.byref(p), None, 0, 10, 100, C.c_void_p(1), None)
    assert rc == _abi.EINVAL and b"incomplete" in lib.dsdx_last_error()
    rc = lib.dsdx_halo_build(plan, C.byref(box), C.byref(p), None, 0, 10, 5, C.c_void_p(1), None)
    assert rc == _abi.EINVAL
    fields = (_abi.Field * 17)()
    rc = lib.dsdx_forward_exchange(plan, fields, 17, 0, C.c_void_p(1))
    assert rc == _abi.EINVAL and b"16" in lib.dsdx_last_error()
    # the allocation-free build checks its fixed limit before any device work
    rc = lib.dsdx_halo_build_noalloc(C.byref(box), C.byref(p), None, 0, 262145, 10 ** 6, None,
                                     C.c_void_p(1), None)
    assert rc == _abi.EINVAL and b"262144" in lib.dsdx_last_error()
    rc = lib.dsdx_halo_build_noalloc(C.byref(box), C.byref(p), None, 0, 10, 10, None, None, None)
    assert rc == _abi.EINVAL and b"stream" in lib.dsdx_last_error()
    # no device halo yet: reverse and forward are no-ops, refresh is a state error
    assert lib.dsdx_reverse_halo(plan, None, None, None, 0, C.c_void_p(1)) == 0
    rc = lib.dsdx_halo_refresh(plan, None, C.byref(p), None, 0, 0, 0, C.c_void_p(1))
    assert rc == _abi.ESTATE
    assert lib.dsdx_plan_destroy(plan) == 0


def test_record_layouts_match_the_oracle(lib):
    """dsdx_wire_layout (host only) against the oracle's restatement of
    make_device_wire_layout and the exchange-record rule, for every field type
    mix below; and the plan-aware record calls refuse a plan without a halo."""
    import numpy as np

    from dsdecompx_amd import _abi
    from oracle import pyoracle as po
    mixes = [(), ((np.float64, 1),), ((np.float32, 3),), ((np.float32, 1), (np.float64, 1)),
             ((np.int32, 1), (np.int64, 2), (np.uint32, 1)), ((np.uint64, 5), (np.float32, 1), (np.int64, 1))]
    code = {np.int32: _abi.INT32, np.int64: _abi.INT64, np.uint32: _abi.UINT32, np.uint64: _abi.UINT64,
            np.float32: _abi.FLOAT32, np.float64: _abi.FLOAT64}
    for mix in mixes:
        ofields = [po.Field(np.zeros((4, c) if c > 1 else 4, dt)) for dt, c in mix]
        arr = (_abi.Field * max(1, len(mix)))()
        for k, (dt, c) in enumerate(mix):
            arr[k] = _abi.Field(1, code[dt], c, 0)
        for kind in (0, 1):
            offs = (C.c_uint32 * max(1, len(mix)))()
            stride = C.c_uint32()
            assert lib.dsdx_wire_layout(arr, len(mix), kind, offs, C.byref(stride)) == 0
            assert ([offs[k] for k in range(len(mix))], stride.value) == po.wire_layout(ofields, kind)
    assert lib.dsdx_wire_layout(None, 0, 2, None, None) == _abi.EINVAL
    plan = C.c_void_p()
    assert lib.dsdx_plan_create(C.byref(plan)) == 0
    p = _abi.Particles()
    rc = lib.dsdx_halo_pack(plan, C.byref(p), None, 0, 0, 0, C.c_void_p(8), C.c_void_p(1))
    assert rc == _abi.ESTATE
    rc = lib.dsdx_loopback_exchange(plan, None, None, 96, 0, C.c_void_p(1))
    assert rc == _abi.ESTATE
    assert lib.dsdx_plan_destroy(plan) == 0
