"""Pins the oracle's restatement of the wire format (oracle/dsdx_oracle_wire.c)
against what fixes it independently of its code: the member offsets of
ParticleRecord (include/field.hpp:150-157) decoded with a numpy record type,
stride known answers worked out by hand from make_device_wire_layout
(src/device_fields.cpp:35-51) and the exchange-record rule
(src/halo.cpp:1253-1271), and round trips through the loopback exchange that
must reproduce the oracle's direct Host results bit for bit."""
import numpy as np
import pytest

from oracle import pyoracle as po

# ParticleRecord as the reference's C++ compiler lays it out: 11 doubles, then
# two 4-byte ints, no padding
RECORD = np.dtype([(n, "<f8") for n in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz",
                                        "mass", "charge")] + [("type", "<i4"), ("global_id", "<i4")])


def _fields(cap, rng, spec):
    out = []
    for dt, comps, op in spec:
        shape = (cap,) if comps == 1 else (cap, comps)
        if np.issubdtype(dt, np.floating):
            a = rng.standard_normal(shape).astype(dt)
        else:
            a = rng.integers(0, 1000, shape).astype(dt)
        out.append(po.Field(a, op))
    return out


def _setup(n=600, L=10.0, w=1.3, seed=3, periodic=(True, True, True), corners=True, spec=()):
    rng = np.random.default_rng(seed)
    box = po.Box(upper=(L, L, L), periodic=periodic, width=w, edges_and_corners=corners)
    ora = po.Oracle(box)
    cap = n * 8
    p = po.Particles(cap)
    p.x[:n], p.y[:n], p.z[:n] = rng.uniform(0, L, (3, n))
    for k in ("vx", "vy", "vz", "fx", "fy", "fz", "charge"):
        getattr(p, k)[:n] = rng.standard_normal(n)
    p.mass[:n] = rng.uniform(0.5, 2, n)
    p.type[:n] = rng.integers(0, 5, n)
    p.global_id[:n] = rng.permutation(n * 3)[:n]
    fields = _fields(cap, rng, spec)
    g = ora.build(p, n)
    return ora, p, n, g, fields, rng


def _copy(p):
    q = po.Particles(p.capacity)
    for k in po.Particles.F64 + po.Particles.I32:
        getattr(q, k)[:] = getattr(p, k)
    return q


def _copy_fields(fields):
    return [po.Field(f.data.copy(), f.op) for f in fields]


def test_record_member_offsets():
    assert RECORD.itemsize == 96 == po.RECORD_BYTES
    box = po.Box(upper=(4.0, 4.0, 4.0), width=0.5)
    ora = po.Oracle(box)
    p = po.Particles(64)
    p.x[0], p.y[0], p.z[0] = 0.1, 2.0, 2.0          # one image, through slot -x
    p.vx[0], p.vy[0], p.vz[0] = 1.5, -2.5, 3.25
    p.fx[0], p.fy[0], p.fz[0] = 9.0, 9.0, 9.0       # forces are not sent
    p.mass[0], p.charge[0], p.type[0], p.global_id[0] = 0.75, -1.0, 7, 123456
    assert ora.build(p, 1) == 1
    buf = po.Wire(ora).pack(p)
    rec = np.frombuffer(buf.tobytes(), RECORD)
    assert len(rec) == 1
    r = rec[0]
    # slot -x: the image sits one extent to the right
    assert (r["x"], r["y"], r["z"]) == (0.1 + 4.0, 2.0, 2.0)
    assert (r["vx"], r["vy"], r["vz"]) == (1.5, -2.5, 3.25)
    assert (r["fx"], r["fy"], r["fz"]) == (0.0, 0.0, 0.0)
    assert (r["mass"], r["charge"], r["type"], r["global_id"]) == (0.75, -1.0, 7, 123456)


@pytest.mark.parametrize("spec,kind,offsets,stride", [
    ((), po.WIRE_PARTICLE, [], 96),
    (((np.float64, 1),), po.WIRE_PARTICLE, [96], 104),
    (((np.float32, 3),), po.WIRE_PARTICLE, [96], 112),                       # 108 -> 112
    (((np.float32, 1), (np.float64, 1)), po.WIRE_PARTICLE, [96, 100], 112),  # no per-field alignment
    (((np.int32, 1), (np.int64, 2), (np.uint32, 1)), po.WIRE_PARTICLE, [96, 100, 116], 120),
    (((np.float64, 1),), po.WIRE_EXCHANGE, [0], 8),
    (((np.float32, 3),), po.WIRE_EXCHANGE, [0], 16),                         # 12 -> 16
    (((np.float32, 1), (np.float64, 1)), po.WIRE_EXCHANGE, [0, 8], 16),      # 4 -> aligned 8
    (((np.int32, 1), (np.int64, 2), (np.uint32, 1)), po.WIRE_EXCHANGE, [0, 8, 24], 32),
    (((np.uint32, 3), (np.float32, 1)), po.WIRE_EXCHANGE, [0, 12], 16),
])
def test_layout_known_answers(spec, kind, offsets, stride):
    fields = [po.Field(np.zeros((4, c) if c > 1 else 4, dt)) for dt, c in spec]
    assert po.wire_layout(fields, kind) == (offsets, stride)


SPEC = ((np.float64, 1, po.SUM), (np.float32, 3, po.MIN), (np.int64, 2, po.MAX),
        (np.uint32, 1, po.REPLACE), (np.int32, 1, po.SUM))


@pytest.mark.parametrize("periodic,corners", [((True, True, True), True),
                                              ((True, True, False), True),
                                              ((True, True, True), False)])
def test_pack_records_decode_to_owner_plus_shift(periodic, corners):
    ora, p, n, g, fields, _ = _setup(periodic=periodic, corners=corners, spec=SPEC)
    offsets, stride = po.wire_layout(fields, po.WIRE_PARTICLE)
    buf = po.Wire(ora).pack(p, fields)
    assert len(buf) == g * stride
    k = 0
    nb = ora.nb
    for m in range(nb.raw.count):
        cnt = int(ora.plan.raw.send_count[m])
        for j in range(cnt):
            i = int(ora.plan.raw.source[m][j])
            rec = np.frombuffer(buf[k * stride:k * stride + 96].tobytes(), RECORD)[0]
            assert rec["x"] == p.x[i] + nb.shift[m][0] and rec["y"] == p.y[i] + nb.shift[m][1]
            assert rec["z"] == p.z[i] + nb.shift[m][2]
            assert rec["global_id"] == p.global_id[i] and rec["fx"] == 0.0
            for f, off in zip(fields, offsets):
                b = f.data[i].tobytes()
                assert buf[k * stride + off:k * stride + off + len(b)].tobytes() == b
            k += 1
    assert k == g


@pytest.mark.parametrize("periodic", [(True, True, True), (True, False, True)])
def test_forward_round_trip_equals_write_ghosts(periodic):
    ora, p, n, g, fields, _ = _setup(periodic=periodic, spec=SPEC)
    ref, ref_f = _copy(p), _copy_fields(fields)
    ora.write_ghosts(ref, n, ref_f)
    wire = po.Wire(ora)
    stride = po.wire_layout(fields, po.WIRE_PARTICLE)[1]
    recv = wire.exchange(wire.pack(p, fields), stride)
    got, got_f = _copy(p), _copy_fields(fields)
    wire.unpack(recv, g, got, n, got_f)
    for k in po.Particles.F64 + po.Particles.I32:
        np.testing.assert_array_equal(getattr(got, k).view(np.uint8), getattr(ref, k).view(np.uint8), k)
    for a, b in zip(got_f, ref_f):
        np.testing.assert_array_equal(a.data.view(np.uint8), b.data.view(np.uint8))


def test_positions_round_trip():
    ora, p, n, g, _, rng = _setup()
    ref = _copy(p)
    ora.write_ghosts(ref, n)
    wire = po.Wire(ora)
    got = _copy(p)
    got.fx[n:n + g] = 5.0
    recv = wire.exchange(wire.pack(p, positions_only=True), 24)
    wire.unpack_positions(recv, g, got, n)
    for k in ("x", "y", "z", "fx", "fy", "fz"):
        np.testing.assert_array_equal(getattr(got, k)[:n + g], getattr(ref, k)[:n + g], k)
    # the other members of the ghosts are left alone
    assert (got.global_id[n:n + g] == p.global_id[n:n + g]).all()


def test_reverse_records_equal_reverse_halo():
    ora, p, n, g, _, rng = _setup(n=900, seed=8)
    ora.write_ghosts(p, n)
    # wide exponent range: any other summation order changes the low bits
    for k in ("fx", "fy", "fz"):
        getattr(p, k)[:n + g] = rng.standard_normal(n + g) * 10.0 ** rng.integers(-8, 9, n + g)
    ref = _copy(p)
    ora.reverse_halo(ref, n)
    wire = po.Wire(ora)
    back = wire.exchange(wire.pack_reverse(p, n, g), po.RECORD_BYTES, reverse=True)
    rec = np.frombuffer(wire.pack_reverse(p, n, g).tobytes(), RECORD)
    assert (rec["x"] == 0).all() and (rec["global_id"] == 0).all()
    np.testing.assert_array_equal(rec["fx"], p.fx[n:n + g])
    wire.accumulate_reverse(back, p)
    for k in ("fx", "fy", "fz"):
        np.testing.assert_array_equal(getattr(p, k)[:n].view(np.uint64), getattr(ref, k)[:n].view(np.uint64))


@pytest.mark.parametrize("op", [po.SUM, po.MIN, po.MAX, po.REPLACE])
def test_field_records_equal_exchanges(op):
    spec = tuple((dt, c, op) for dt, c, _ in SPEC) + ((np.uint64, 1, op), (np.float64, 3, op))
    ora, p, n, g, fields, rng = _setup(spec=spec, seed=11)
    wire = po.Wire(ora)
    stride = po.wire_layout(fields, po.WIRE_EXCHANGE)[1]
    # forward: owner values into the ghosts
    ref = _copy_fields(fields)
    ora.forward_exchange(n, ref)
    got = _copy_fields(fields)
    wire.scatter_fields(got, wire.exchange(wire.gather_fields(got), stride), g, n)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a.data.view(np.uint8), b.data.view(np.uint8))
    # reverse: ghost values reduced into their owners
    ref = _copy_fields(got)
    ora.reverse_exchange(n, ref)
    back = wire.exchange(wire.pack_fields(got, n, g), stride, reverse=True)
    wire.reduce_fields(got, back)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a.data[:n].view(np.uint8), b.data[:n].view(np.uint8))


def test_empty_plan():
    ora, p, n, g, fields, _ = _setup(n=4, L=40.0, w=0.1, seed=2, spec=SPEC)
    p.x[:n] = p.y[:n] = p.z[:n] = 20.0
    assert ora.build(p, n) == 0
    wire = po.Wire(ora)
    assert len(wire.pack(p, fields)) == 0 and len(wire.gather_fields(fields)) == 0
