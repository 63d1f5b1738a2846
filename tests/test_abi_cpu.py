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
    # no device halo yet: reverse and forward are no-ops, refresh is a state error
    assert lib.dsdx_reverse_halo(plan, None, None, None, 0, C.c_void_p(1)) == 0
    rc = lib.dsdx_halo_refresh(plan, None, C.byref(p), None, 0, 0, 0, C.c_void_p(1))
    assert rc == _abi.ESTATE
    assert lib.dsdx_plan_destroy(plan) == 0
