"""Communication records on the GPU (dsdx_halo_pack and the other record
entry points): every packed buffer byte-identical with the oracle's
restatement of the reference's wire format (oracle/dsdx_oracle_wire.c), and
pack -> loopback exchange -> unpack / accumulate / reduce equal, bit for bit,
to the direct device path (which is itself bit-exact with the Host path)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import F64, I32, assert_bits_equal, host, say, synth, to_device, to_oracle

from tests.gpu_util import box_ghost_writes  # noqa: F401  (autouse: every ghost-write kernel)
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


CASES = [
    # name, particles, L, width, periodic, edges_and_corners
    ("periodic", 6000, 14.0, 2.0, (True, True, True), True),
    ("z_walls", 5000, 13.0, 1.7, (True, True, False), True),
    ("faces_only", 4000, 12.0, 1.5, (True, True, True), False),
]
# (numpy dtype, torch dtype name, components, op)
FIELDS = [(np.float64, "float64", 1, po.SUM), (np.float32, "float32", 3, po.MIN),
          (np.int64, "int64", 2, po.MAX), (np.uint32, "uint32", 1, po.REPLACE),
          (np.int32, "int32", 1, po.SUM)]


def _setup(torch, case, seed=5):
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    name, n, L, w, periodic, corners = case
    d = synth(n, L, seed)
    cap = 8 * n
    ora = po.Oracle(po.Box(upper=(L,) * 3, periodic=periodic, width=w, edges_and_corners=corners))
    op = to_oracle(d, cap)
    dev = to_device(d, cap)
    rng = np.random.default_rng(seed)
    of, df = [], []
    for dt, tname, comps, rop in FIELDS:
        shape = (cap,) if comps == 1 else (cap, comps)
        a = (rng.standard_normal(shape) if np.issubdtype(dt, np.floating)
             else rng.integers(0, 1000, shape)).astype(dt)
        of.append(po.Field(a.copy(), rop))
        df.append(torch.from_numpy(a.copy()).cuda())
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, periodic=periodic, width=w, edges_and_corners=corners))
    return ora, op, of, halo, dev, df, n


def _bytes(t):
    return host(t).tobytes()


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_records_match_oracle_and_direct_path(gpu, case):
    torch = gpu
    from dsdecompx_amd.halo import DeviceParticles
    from dsdecompx_amd.wire import PARTICLE, Records, layout
    say("wire", case[0])
    ora, op, of, halo, dev, df, n = _setup(torch, case)
    g = ora.build(op, n)
    wire_o = po.Wire(ora)
    assert halo.exchange_halo(dev, n) == g        # plan only; the ghosts come from the records
    rec = Records(halo)
    assert layout(df, PARTICLE) == po.wire_layout(of, po.WIRE_PARTICLE) and layout(df, PARTICLE)[1] > 96

    # forward particle records: byte-identical send buffer
    send = rec.pack(dev, n, df)
    expect = wire_o.pack(op, of)
    assert _bytes(send) == expect.tobytes()
    # loopback, then unpack into a fresh copy: the direct build's ghosts
    stride = layout(df, PARTICLE)[1]
    recv = rec.loopback(send, stride)
    assert _bytes(recv) == wire_o.exchange(expect, stride).tobytes()
    got = DeviceParticles(dev.capacity)
    for k in F64 + I32:
        getattr(got, k)[:n] = getattr(dev, k)[:n]
    got_f = [f.clone() for f in df]
    rec.unpack(recv, g, got, n, got_f)
    ref_f = [f.clone() for f in df]
    halo.exchange_halo(dev, n, ref_f)              # the direct path, with fields
    for k in F64 + I32:
        assert_bits_equal(host(getattr(got, k))[:n + g], host(getattr(dev, k))[:n + g], f"unpack {k}")
    for a, b in zip(got_f, ref_f):
        assert _bytes(a[:n + g]) == _bytes(b[:n + g])

    # position records
    pos = rec.pack(dev, n, positions_only=True)
    assert _bytes(pos) == wire_o.pack(op, positions_only=True).tobytes()
    got.fx.fill_(7.0)
    rec.unpack_positions(rec.loopback(pos, 24), g, got, n)
    for k in ("x", "y", "z"):
        assert_bits_equal(host(getattr(got, k))[:n + g], host(getattr(dev, k))[:n + g], f"positions {k}")
    assert (host(got.fx)[n:n + g] == 0.0).all()

    # returned force records: accumulate == reverse_halo
    rng = np.random.default_rng(17)
    for k in ("fx", "fy", "fz"):
        v = rng.standard_normal(n + g) * 10.0 ** rng.integers(-8, 9, n + g)
        getattr(dev, k)[:n + g] = torch.from_numpy(v).cuda()
        getattr(got, k)[:n + g] = torch.from_numpy(v).cuda()
        getattr(op, k)[:n + g] = v
    back = rec.pack_reverse(got, n, g)
    assert _bytes(back) == wire_o.pack_reverse(op, n, g).tobytes()
    rec.accumulate_reverse(rec.loopback(back, 96, reverse=True), got, n)
    halo.reverse_halo(dev, n)
    ora.reverse_halo(op, n)
    for k in ("fx", "fy", "fz"):
        assert_bits_equal(host(getattr(got, k))[:n], host(getattr(dev, k))[:n], f"accumulate {k}")
        assert_bits_equal(host(getattr(got, k))[:n], getattr(op, k)[:n], f"accumulate vs oracle {k}")


@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: c[0])
def test_field_records_match_exchanges(gpu, case):
    torch = gpu
    from dsdecompx_amd.wire import EXCHANGE, Records, layout
    say("wire fields", case[0])
    ora, op, of, halo, dev, df, n = _setup(torch, case, seed=9)
    g = ora.build(op, n)
    wire_o = po.Wire(ora)
    assert halo.exchange_halo(dev, n) == g
    rec = Records(halo)
    offsets, stride = layout(df, EXCHANGE)
    assert (offsets, stride) == po.wire_layout(of, po.WIRE_EXCHANGE)
    # forward: gathered owner records, byte-identical; scatter == forward_exchange
    sent = rec.gather_fields(df, n)
    assert _bytes(sent) == wire_o.gather_fields(of).tobytes()
    got = [f.clone() for f in df]
    rec.scatter_fields(got, rec.loopback(sent, stride), g, n)
    ref = [f.clone() for f in df]
    halo.forward_exchange(ref, n)
    for a, b in zip(got, ref):
        assert _bytes(a[:n + g]) == _bytes(b[:n + g])
    # reverse: fresh ghost values back to their owners, reduced with each op
    rng = np.random.default_rng(23)
    for t, (dt, _tn, comps, _op) in zip(got, FIELDS):
        shape = (g,) if comps == 1 else (g, comps)
        v = (rng.standard_normal(shape) if np.issubdtype(dt, np.floating)
             else rng.integers(0, 1000, shape)).astype(dt)
        t[n:n + g] = torch.from_numpy(v).cuda()
    packed = rec.pack_fields(got, n, g)
    assert _bytes(packed) == wire_o.pack_fields([po.Field(host(t).copy(), f.op) for t, f in zip(got, of)],
                                               n, g).tobytes()
    red = [f.clone() for f in got]
    rec.reduce_fields(list(zip(red, [f.op for f in of])), rec.loopback(packed, stride, reverse=True), n)
    direct = [f.clone() for f in got]
    halo.reverse_exchange(list(zip(direct, [f.op for f in of])), n)
    for a, b in zip(red, direct):
        assert _bytes(a[:n]) == _bytes(b[:n])


def test_empty_halo_records(gpu):
    torch = gpu
    from dsdecompx_amd.halo import BoxConfig, DeviceParticles, LoopbackHalo
    from dsdecompx_amd.wire import Records
    say("wire empty")
    dev = DeviceParticles(16)
    dev.x[:4] = 20.0
    dev.y[:4] = 20.0
    dev.z[:4] = 20.0
    halo = LoopbackHalo(BoxConfig(upper=(40.0,) * 3, width=0.5))
    assert halo.exchange_halo(dev, 4) == 0
    rec = Records(halo)
    send = rec.pack(dev, 4)
    assert send.numel() == 0
    rec.unpack(rec.loopback(send, 96), 0, dev, 4)
    rec.accumulate_reverse(rec.loopback(rec.pack_reverse(dev, 4, 0), 96, reverse=True), dev, 4)
    torch.cuda.synchronize()


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_build_records_equals_oracle_pack(gpu, case):
    """dsdx_halo_build_records (classify, route, pack from the classification
    pass's positions, no ghost written): the oracle's send buffer byte for
    byte; the particle arrays past `owned` untouched; a send buffer one record
    short gives DSDX_ECAPACITY, after which dsdx_halo_pack fills a larger one."""
    torch = gpu
    import ctypes as C
    from dsdecompx_amd import _abi
    from dsdecompx_amd.wire import PARTICLE, Records, _buffer, _ptr, layout
    say("build records", case[0])
    ora, op, of, halo, dev, df, n = _setup(torch, case, seed=13)
    g = ora.build(op, n)
    expect = po.Wire(ora).pack(op, of).tobytes()
    rec = Records(halo)
    dev.x[n:].fill_(-3.0)
    send, got = rec.build(dev, n, df, capacity=g + 5)
    assert got == g and _bytes(send) == expect
    assert (host(dev.x)[n:] == -3.0).all()
    # one record short
    stride = layout(df, PARTICLE)[1]
    small = _buffer(max(0, g - 1) * stride, dev.x.device)
    arr = (_abi.Field * len(df))()
    from dsdecompx_amd.halo import field_struct
    for k, t in enumerate(df):
        arr[k] = field_struct(t)
    cnt = C.c_uint64()
    pc = dev.c()
    rc = _abi.lib().dsdx_halo_build_records(halo.plan, C.byref(halo._box), C.byref(pc), arr, len(df), n,
                                            _ptr(small), max(0, g - 1),
                                            torch.cuda.current_stream().cuda_stream, C.byref(cnt))
    if g:
        assert rc == _abi.ECAPACITY and cnt.value == g
        assert _bytes(rec.pack(dev, n, df)) == expect
    # the records feed the ghosts of the direct path
    recv = rec.loopback(send, stride)
    rec.unpack(recv, g, dev, n, df)
    ref = to_device(synth(*case[1:3], 13), dev.capacity)
    ref_f = [f.clone() for f in df]
    halo.exchange_halo(ref, n, ref_f)
    for k in F64 + I32:
        assert_bits_equal(host(getattr(dev, k))[:n + g], host(getattr(ref, k))[:n + g], f"ghosts {k}")


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_reverse_halo_after_record_build(gpu, case):
    """dsdx_reverse_halo on a plan built through the record path, whose ghosts
    were unpacked from its records: that build writes the owner maps (its
    route pass) and no chunk rows, so reverse_halo takes the owner-map
    kernels whatever the ghost-write mode, and the sums match the oracle."""
    torch = gpu
    from dsdecompx_amd.wire import PARTICLE, Records, layout
    say("reverse after record build", case[0])
    ora, op, of, halo, dev, df, n = _setup(torch, case, seed=21)
    g = ora.build(op, n)
    rec = Records(halo)
    send, got = rec.build(dev, n, df, capacity=g + 5)
    assert got == g
    rec.unpack(rec.loopback(send, layout(df, PARTICLE)[1]), g, dev, n, df)
    rng = np.random.default_rng(23)
    for k in ("fx", "fy", "fz"):
        v = rng.uniform(-1.0, 1.0, n + g)
        getattr(dev, k)[:n + g] = torch.from_numpy(v).cuda()
        getattr(op, k)[:n + g] = v
    halo.reverse_halo(dev, n)
    ora.reverse_halo(op, n)
    for k in ("fx", "fy", "fz"):
        assert_bits_equal(host(getattr(dev, k))[:n], getattr(op, k)[:n], f"reverse {k}")
