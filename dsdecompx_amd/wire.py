"""Communication records of a single-box halo plan (``dsdx_halo_pack`` and
the other record entry points of include/dsdx_hip.h): the reference's wire
formats -- ParticleRecord (include/field.hpp:150-157) plus the registered
fields, exchange records of registered fields, position records -- for
callers that carry them over a transport of their own.  Buffers are torch
uint8 device tensors; ``loopback`` moves them as the self messages of the
1x1x1 grid do."""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .halo import DeviceParticles, LoopbackHalo, _field_array, _stream_ptr, field_struct

RECORD_BYTES = 96
PARTICLE, EXCHANGE = 0, 1
POSITION_BYTES = 24


def layout(fields, kind: int):
    """(byte offset of each field, record stride) of a particle record
    (kind PARTICLE) or an exchange record (kind EXCHANGE)."""
    arr, n = _field_array(fields)
    offs = (C.c_uint32 * max(1, n))()
    stride = C.c_uint32()
    _abi.check(_abi.lib().dsdx_wire_layout(arr, n, kind, offs, C.byref(stride)), "wire layout")
    return [int(offs[f]) for f in range(n)], int(stride.value)


def _typed(fields):
    """[(tensor, op)] -> field table with the ops."""
    arr = (_abi.Field * max(1, len(fields)))()
    for k, (t, op) in enumerate(fields):
        arr[k] = field_struct(t, op)
    return arr, len(fields)


def _ptr(t: torch.Tensor) -> int:
    """Device address of a buffer, also of an empty one (torch reports 0 for
    empty tensors; the entry points want a valid address)."""
    return t.data_ptr() if t.numel() else t.untyped_storage().data_ptr()


def _buffer(nbytes: int, device) -> torch.Tensor:
    # 8-byte aligned storage, viewed as bytes
    words = (max(1, nbytes) + 7) // 8
    return torch.zeros(words, dtype=torch.float64, device=device).view(torch.uint8)[:nbytes]


class Records:
    """The record steps of one ``LoopbackHalo``'s last build."""

    def __init__(self, halo: LoopbackHalo):
        self.halo = halo

    def build(self, p: DeviceParticles, owned: int, fields=(), capacity=None, stream=None):
        """dsdx_halo_build_records: classify, route and pack the particle
        records of a new halo (no ghost written to `p`).  Returns
        (send buffer, records); grows the buffer once if `capacity` (records)
        is too small."""
        stride = layout(fields, PARTICLE)[1]
        cap = int(capacity if capacity is not None else max(1024, owned // 8))
        out = _buffer(cap * stride, p.x.device)
        arr, n = _field_array(fields)
        pc = p.c()
        box = self.halo._box
        g = C.c_uint64()
        rc = _abi.lib().dsdx_halo_build_records(self.halo.plan, C.byref(box), C.byref(pc), arr, n, owned,
                                                _ptr(out), cap, _stream_ptr(stream), C.byref(g))
        if rc == _abi.ECAPACITY:
            out = self.pack(p, owned, fields, stream=stream)
        else:
            _abi.check(rc, "halo build records")
            out = out[:g.value * stride]
        self.halo.owned = owned
        self.halo.ghosts = int(g.value)
        return out, int(g.value)

    def counts(self, stream=None):
        """(records sent, records received): the plan's send and ghost totals."""
        send, recv, g, _, _ = self.halo.query(stream)
        return sum(send), g

    def pack(self, p: DeviceParticles, owned: int, fields=(), positions_only=False, stream=None):
        sent, _ = self.counts(stream)
        stride = POSITION_BYTES if positions_only else layout(fields, PARTICLE)[1]
        out = _buffer(sent * stride, p.x.device)
        arr, n = _field_array(() if positions_only else fields)
        pc = p.c()
        _abi.check(_abi.lib().dsdx_halo_pack(self.halo.plan, C.byref(pc), arr, n, owned,
                                             1 if positions_only else 0, _ptr(out),
                                             _stream_ptr(stream)), "halo pack")
        return out

    def loopback(self, send: torch.Tensor, record_bytes: int, reverse=False, stream=None):
        out = _buffer(send.numel(), send.device)
        _abi.check(_abi.lib().dsdx_loopback_exchange(self.halo.plan, _ptr(send), _ptr(out),
                                                     record_bytes, 1 if reverse else 0,
                                                     _stream_ptr(stream)), "loopback exchange")
        return out

    @staticmethod
    def unpack(records: torch.Tensor, count: int, p: DeviceParticles, dest: int, fields=(), stream=None):
        arr, n = _field_array(fields)
        pc = p.c()
        _abi.check(_abi.lib().dsdx_unpack_records(_ptr(records), count, C.byref(pc), dest, arr, n,
                                                  _stream_ptr(stream)), "unpack records")

    @staticmethod
    def unpack_positions(records: torch.Tensor, count: int, p: DeviceParticles, dest: int, stream=None):
        pc = p.c()
        _abi.check(_abi.lib().dsdx_unpack_positions(_ptr(records), count, C.byref(pc), dest,
                                                    _stream_ptr(stream)), "unpack positions")

    @staticmethod
    def pack_reverse(p: DeviceParticles, first: int, count: int, stream=None):
        out = _buffer(count * RECORD_BYTES, p.fx.device)
        _abi.check(_abi.lib().dsdx_pack_reverse_records(p.fx.data_ptr(), p.fy.data_ptr(), p.fz.data_ptr(),
                                                        first, count, _ptr(out),
                                                        _stream_ptr(stream)), "pack reverse")
        return out

    def accumulate_reverse(self, records: torch.Tensor, p: DeviceParticles, owned: int, stream=None):
        _abi.check(_abi.lib().dsdx_accumulate_reverse_records(
            self.halo.plan, _ptr(records), p.fx.data_ptr(), p.fy.data_ptr(), p.fz.data_ptr(),
            owned, _stream_ptr(stream)), "accumulate reverse")

    @staticmethod
    def pack_fields(fields, first: int, count: int, stream=None):
        stride = layout(fields, EXCHANGE)[1]
        out = _buffer(count * stride, fields[0].device)
        arr, n = _field_array(fields)
        _abi.check(_abi.lib().dsdx_pack_field_records(arr, n, first, count, _ptr(out),
                                                      _stream_ptr(stream)), "pack field records")
        return out

    @staticmethod
    def scatter_fields(fields, records: torch.Tensor, count: int, first: int, stream=None):
        arr, n = _field_array(fields)
        _abi.check(_abi.lib().dsdx_scatter_field_records(arr, n, _ptr(records), count, first,
                                                         _stream_ptr(stream)), "scatter field records")

    def gather_fields(self, fields, owned: int, stream=None):
        sent, _ = self.counts(stream)
        stride = layout(fields, EXCHANGE)[1]
        out = _buffer(sent * stride, fields[0].device)
        arr, n = _field_array(fields)
        _abi.check(_abi.lib().dsdx_gather_field_records(self.halo.plan, arr, n, owned, _ptr(out),
                                                        _stream_ptr(stream)), "gather field records")
        return out

    def reduce_fields(self, fields, records: torch.Tensor, owned: int, stream=None):
        """fields: list of (tensor, op)."""
        arr, n = _typed(fields)
        _abi.check(_abi.lib().dsdx_reduce_field_records(self.halo.plan, arr, n, _ptr(records), owned,
                                                        _stream_ptr(stream)), "reduce field records")
