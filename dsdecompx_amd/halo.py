"""Python host-side mirror of the drop-in path over torch device tensors.

``DeviceParticles`` is the SoA storage of ``ParticleInfo`` on the GPU
(include/dsdecompx.hpp:70-168): 11 float64 and 2 int32 arrays of length
``capacity``, owned particles first, ghosts after them.  ``LoopbackHalo``
holds one device halo plan and calls the C ABI for migrate / exchange_halo /
reverse_halo / forward_exchange / reverse_exchange, with the reference's
argument meaning and errors (src/runtime.cpp:862-1071).

torch only provides device memory and the stream; every kernel is in the
HIP library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _abi

F64_NAMES = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "charge")
I32_NAMES = ("type", "global_id")
NAMES = F64_NAMES + I32_NAMES

_FIELD_TYPE = {torch.int32: _abi.INT32, torch.int64: _abi.INT64, torch.uint32: _abi.UINT32,
               torch.uint64: _abi.UINT64, torch.float32: _abi.FLOAT32,
               torch.float64: _abi.FLOAT64}


def _stream_ptr(stream) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


@dataclass
class BoxConfig:
    """GlobalDomain + HaloConfig (include/domain.hpp:11-30).  `local` =
    (lower, upper) of the local domain whose faces the slab tests use; None:
    the cuts of a fresh 1x1x1 topology, lower + extent * i / 1
    (src/topology.cpp:66-72)."""
    lower: tuple = (0.0, 0.0, 0.0)
    upper: tuple = (1.0, 1.0, 1.0)
    periodic: tuple = (True, True, True)
    width: float = 0.0
    edges_and_corners: bool = True
    local: tuple = None

    def c(self) -> _abi.Box:
        b = _abi.Box()
        for a in range(3):
            b.lower[a] = float(self.lower[a])
            b.upper[a] = float(self.upper[a])
            b.periodic[a] = 1 if self.periodic[a] else 0
        b.width = float(self.width)
        b.edges_and_corners = 1 if self.edges_and_corners else 0
        if self.local is not None:
            for a in range(3):
                b.local_lower[a] = float(self.local[0][a])
                b.local_upper[a] = float(self.local[1][a])
            b.local_given = 1
        return b


class DeviceParticles:
    """Library-owned device SoA storage (ParticleInfo::allocate)."""

    def __init__(self, capacity: int, device="cuda"):
        self.capacity = int(capacity)
        self.device = torch.device(device)
        for n in F64_NAMES:
            setattr(self, n, torch.zeros(self.capacity, dtype=torch.float64, device=self.device))
        for n in I32_NAMES:
            setattr(self, n, torch.zeros(self.capacity, dtype=torch.int32, device=self.device))

    def grow(self, capacity: int, keep: int) -> bool:
        """resize_storage (src/runtime.cpp:688-737): at least 1.5x, keeps the
        first `keep` particles.  Returns True when the arrays moved."""
        if capacity <= self.capacity:
            return False
        capacity = max(capacity, 64, self.capacity + self.capacity // 2)
        for n in NAMES:
            old = getattr(self, n)
            new = torch.zeros(capacity, dtype=old.dtype, device=self.device)
            new[:keep] = old[:keep]
            setattr(self, n, new)
        self.capacity = capacity
        return True

    def c(self) -> _abi.Particles:
        p = _abi.Particles()
        for n in NAMES:
            t = getattr(self, n)
            setattr(p, n, t.data_ptr() if t is not None else None)
        return p


def field_struct(tensor: torch.Tensor, op: int = _abi.SUM) -> _abi.Field:
    if not tensor.is_contiguous():
        raise ValueError("fields must be contiguous")
    comps = 1 if tensor.dim() == 1 else int(tensor[0].numel())
    return _abi.Field(tensor.data_ptr(), _FIELD_TYPE[tensor.dtype], comps, op)


def _field_array(fields):
    fields = list(fields or [])
    arr = (_abi.Field * max(1, len(fields)))()
    for k, f in enumerate(fields):
        arr[k] = f if isinstance(f, _abi.Field) else field_struct(f)
    return arr, len(fields)


class LoopbackHalo:
    """One box on a 1x1x1 self-loopback grid with its device halo plan."""

    def __init__(self, box: BoxConfig):
        self.box = box
        self._box = box.c()
        self._plan = C.c_void_p()
        _abi.check(_abi.lib().dsdx_plan_create(C.byref(self._plan)), "plan create")
        self.owned = 0
        self.ghosts = 0

    def close(self):
        if self._plan:
            _abi.lib().dsdx_plan_destroy(self._plan)
            self._plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def plan(self):
        return self._plan

    def reset(self) -> None:
        _abi.check(_abi.lib().dsdx_plan_reset(self._plan))
        self.ghosts = 0

    def reserve(self, owned: int, ghosts: int) -> None:
        _abi.check(_abi.lib().dsdx_plan_reserve(self._plan, owned, ghosts))

    def migrate(self, p: DeviceParticles, owned: int, stream=None) -> None:
        _abi.check(_abi.lib().dsdx_migrate_wrap(C.byref(self._box), p.x.data_ptr(),
                                                p.y.data_ptr(), p.z.data_ptr(), owned,
                                                _stream_ptr(stream)), "migrate")
        self.reset()
        self.owned = owned

    def exchange_halo(self, p: DeviceParticles, owned: int, fields=(), stream=None,
                      wait: bool = True) -> int:
        """Rebuild; grows `p` (and the fields are the caller's) when needed."""
        arr, n = _field_array(fields)
        g = C.c_uint64(0)
        s = _stream_ptr(stream)
        pc = p.c()
        # the ghosts must fit the particle arrays and every field
        capacity = min([p.capacity] + [int(f.shape[0]) for f in fields or ()])
        rc = _abi.lib().dsdx_halo_build(self._plan, C.byref(self._box), C.byref(pc), arr, n,
                                        owned, capacity, s, C.byref(g) if wait else None)
        if rc == _abi.ECAPACITY:
            if any(int(f.shape[0]) < owned + g.value for f in fields or ()):
                raise _abi.CapacityError("a registered field is too small for the ghosts")
            p.grow(owned + g.value, owned)
            pc = p.c()
            rc = _abi.lib().dsdx_halo_scatter(self._plan, C.byref(pc), arr, n, owned,
                                              p.capacity, s)
        _abi.check(rc, "exchange_halo")
        self.owned = owned
        self.ghosts = int(g.value) if wait else -1
        return self.ghosts

    def migrate_exchange_halo(self, p: DeviceParticles, owned: int, fields=(), stream=None) -> int:
        """migrate then exchange_halo(rebuild), fused into one pass over x, y, z."""
        arr, n = _field_array(fields)
        g = C.c_uint64(0)
        s = _stream_ptr(stream)
        pc = p.c()
        capacity = min([p.capacity] + [int(f.shape[0]) for f in fields or ()])
        rc = _abi.lib().dsdx_migrate_halo_build(self._plan, C.byref(self._box), C.byref(pc), arr,
                                                n, owned, capacity, s, C.byref(g))
        if rc == _abi.ECAPACITY:
            if any(int(f.shape[0]) < owned + g.value for f in fields or ()):
                raise _abi.CapacityError("a registered field is too small for the ghosts")
            p.grow(owned + g.value, owned)
            pc = p.c()
            rc = _abi.lib().dsdx_halo_scatter(self._plan, C.byref(pc), arr, n, owned,
                                              p.capacity, s)
        _abi.check(rc, "migrate_exchange_halo")
        self.owned = owned
        self.ghosts = int(g.value)
        return self.ghosts

    def refresh(self, p: DeviceParticles, owned: int, fields=(), positions_only=False,
                stream=None) -> None:
        arr, n = _field_array(fields)
        pc = p.c()
        _abi.check(_abi.lib().dsdx_halo_refresh(self._plan, None, C.byref(pc), arr, n, owned,
                                                1 if positions_only else 0,
                                                _stream_ptr(stream)), "refresh")

    def query(self, stream=None):
        send = (C.c_uint32 * 26)()
        recv = (C.c_uint32 * 26)()
        g, u, ov = C.c_uint64(), C.c_uint64(), C.c_int32()
        _abi.check(_abi.lib().dsdx_plan_query(self._plan, send, recv, C.byref(g), C.byref(u),
                                              C.byref(ov), _stream_ptr(stream)))
        return list(send), list(recv), g.value, u.value, ov.value

    def ghost_owners(self, ghosts: int, device="cuda", stream=None) -> torch.Tensor:
        out = torch.empty(max(1, ghosts), dtype=torch.int32, device=device)
        _abi.check(_abi.lib().dsdx_plan_ghost_owners(self._plan, out.data_ptr(),
                                                     _stream_ptr(stream)))
        return out[:ghosts]

    def reverse_halo(self, p: DeviceParticles, owned: int, stream=None) -> None:
        _abi.check(_abi.lib().dsdx_reverse_halo(self._plan, p.fx.data_ptr(), p.fy.data_ptr(),
                                                p.fz.data_ptr(), owned, _stream_ptr(stream)),
                   "reverse_halo")

    def reverse_exchange(self, fields, owned: int, stream=None) -> None:
        """fields: list of (tensor, op)."""
        arr = (_abi.Field * max(1, len(fields)))()
        for k, (t, op) in enumerate(fields):
            arr[k] = field_struct(t, op)
        _abi.check(_abi.lib().dsdx_reverse_exchange(self._plan, arr, len(fields), owned,
                                                    _stream_ptr(stream)), "reverse_exchange")

    def forward_exchange(self, fields, owned: int, stream=None) -> None:
        arr, n = _field_array(fields)
        _abi.check(_abi.lib().dsdx_forward_exchange(self._plan, arr, n, owned,
                                                    _stream_ptr(stream)), "forward_exchange")
