"""Batched replicas: many independent periodic boxes in one set of device SoA
arrays, each phase of the step (migrate, exchange_halo, reverse_halo, field
exchanges) one kernel launch over all of them, and the whole step capturable
in a HIP graph through the library (dsdx_graph_begin / dsdx_graph_end,
``capture_step``).  torch tensors here are only the device arrays and streams.

``ReplicaBatch`` lays the replicas out back to back: replica r owns particle
slots [offset_r, offset_r + capacity_r), owned first.  The per-replica
results are bit-identical to running the single-box path on each replica's
own slots (the plain per-box loop, ``loop_step``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _abi
from .halo import NAMES, BoxConfig, LoopbackHalo, field_struct

# Expected ghost fraction of a full shell of width w around a periodic cube
# of side L: ((L + 2w)^3 - L^3) / L^3; capacity adds a margin on top.
def shell_fraction(L: float, w: float) -> float:
    return ((L + 2.0 * w) ** 3 - L ** 3) / L ** 3


@dataclass
class ReplicaSpec:
    owned: int
    box: BoxConfig


class ReplicaBatch:
    """Device storage + batched halo plan for a list of replicas."""

    def __init__(self, specs, device="cuda", margin: float = 1.35, extra: int = 256,
                 stream=None, capacities=None):
        """`capacities` (optional, one per replica) replaces the shell-based
        estimate: particle slots for the replica's owned particles and ghosts."""
        self.device = torch.device(device)
        self.specs = list(specs)
        self.offsets, self.capacities = [], []
        off = 0
        for r, s in enumerate(self.specs):
            L = max(s.box.upper[a] - s.box.lower[a] for a in range(3))
            cap = s.owned + int(margin * shell_fraction(L, s.box.width) * s.owned) + extra
            if capacities is not None:
                cap = int(capacities[r])
            self.offsets.append(off)
            self.capacities.append(cap)
            off += cap
        self.total = off
        for n in NAMES:
            dt = torch.int32 if n in ("type", "global_id") else torch.float64
            setattr(self, n, torch.zeros(self.total, dtype=dt, device=self.device))
        table = (_abi.Replica * len(self.specs))()
        for r, s in enumerate(self.specs):
            table[r].box = s.box.c()
            table[r].offset = self.offsets[r]
            table[r].capacity = self.capacities[r]
            table[r].owned = s.owned
        self._table = table
        self._batch = C.c_void_p()
        _abi.check(_abi.lib().dsdx_batch_create(table, len(self.specs), self._stream(stream),
                                                C.byref(self._batch)), "batch create")

    @staticmethod
    def _stream(stream):
        if stream is None:
            stream = torch.cuda.current_stream()
        return stream.cuda_stream

    def close(self):
        if self._batch:
            _abi.lib().dsdx_batch_destroy(self._batch)
            self._batch = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._batch

    def particles(self) -> _abi.Particles:
        p = _abi.Particles()
        for n in NAMES:
            setattr(p, n, getattr(self, n).data_ptr())
        return p

    def owned_slice(self, r: int) -> slice:
        return slice(self.offsets[r], self.offsets[r] + self.specs[r].owned)

    def replica_slice(self, r: int, ghosts: int) -> slice:
        return slice(self.offsets[r], self.offsets[r] + self.specs[r].owned + ghosts)

    def set_shape(self, items: int, waves: int, scan_threads: int, stream=None) -> None:
        _abi.check(_abi.lib().dsdx_batch_set_shape(self._batch, items, waves, scan_threads,
                                                   self._stream(stream)), "set shape")

    def shape(self):
        i, w, s = C.c_int32(), C.c_int32(), C.c_int32()
        _abi.check(_abi.lib().dsdx_batch_get_shape(self._batch, C.byref(i), C.byref(w),
                                                   C.byref(s)))
        return i.value, w.value, s.value

    # --- one launch per phase over all replicas --------------------------------
    def migrate(self, stream=None) -> None:
        _abi.check(_abi.lib().dsdx_batch_migrate_wrap(self._batch, self.x.data_ptr(),
                                                      self.y.data_ptr(), self.z.data_ptr(),
                                                      self._stream(stream)), "batch migrate")

    def exchange_halo(self, fields=(), stream=None) -> None:
        arr = (_abi.Field * max(1, len(fields)))()
        for k, f in enumerate(fields):
            arr[k] = field_struct(f)
        p = self.particles()
        _abi.check(_abi.lib().dsdx_batch_halo_build(self._batch, C.byref(p), arr, len(fields),
                                                    self._stream(stream)), "batch halo")

    def migrate_exchange_halo(self, fields=(), stream=None) -> None:
        """migrate then exchange_halo, fused into one pass over x, y, z."""
        arr = (_abi.Field * max(1, len(fields)))()
        for k, f in enumerate(fields):
            arr[k] = field_struct(f)
        p = self.particles()
        _abi.check(_abi.lib().dsdx_batch_migrate_halo_build(self._batch, C.byref(p), arr,
                                                            len(fields), self._stream(stream)),
                   "batch migrate + halo")

    def refresh(self, positions_only=True, stream=None) -> None:
        p = self.particles()
        _abi.check(_abi.lib().dsdx_batch_refresh(self._batch, C.byref(p), None, 0,
                                                 1 if positions_only else 0,
                                                 self._stream(stream)), "batch refresh")

    def reverse_halo(self, stream=None) -> None:
        _abi.check(_abi.lib().dsdx_batch_reverse_halo(self._batch, self.fx.data_ptr(),
                                                      self.fy.data_ptr(), self.fz.data_ptr(),
                                                      self._stream(stream)), "batch reverse")

    def reverse_exchange(self, fields, stream=None) -> None:
        arr = (_abi.Field * max(1, len(fields)))()
        for k, (t, op) in enumerate(fields):
            arr[k] = field_struct(t, op)
        _abi.check(_abi.lib().dsdx_batch_reverse_exchange(self._batch, arr, len(fields),
                                                          self._stream(stream)), "batch reverse")

    def forward_exchange(self, fields, stream=None) -> None:
        arr = (_abi.Field * max(1, len(fields)))()
        for k, t in enumerate(fields):
            arr[k] = field_struct(t)
        _abi.check(_abi.lib().dsdx_batch_forward_exchange(self._batch, arr, len(fields),
                                                          self._stream(stream)), "batch forward")

    def counts(self, stream=None):
        """Waits; (ghosts, owners, status) per replica."""
        n = len(self.specs)
        g, u, s = (C.c_uint32 * n)(), (C.c_uint32 * n)(), (C.c_uint32 * n)()
        _abi.check(_abi.lib().dsdx_batch_counts(self._batch, g, u, s, self._stream(stream)))
        return list(g), list(u), list(s)

    def ghost_owners(self, stream=None) -> torch.Tensor:
        out = torch.full((self.total,), -1, dtype=torch.int32, device=self.device)
        _abi.check(_abi.lib().dsdx_batch_ghost_owners(self._batch, out.data_ptr(),
                                                      self._stream(stream)))
        return out


class LoopedReplicas:
    """The plain per-box loop over the same storage: one single-box plan per
    replica, each phase one C-ABI call per replica (with the ghost-count wait
    of every exchange_halo)."""

    def __init__(self, batch: ReplicaBatch):
        self.batch = batch
        self.halos = [LoopbackHalo(s.box) for s in batch.specs]
        self.ghosts = [0] * len(batch.specs)

    def _view(self, r):
        b = self.batch
        p = _abi.Particles()
        off = b.offsets[r]
        for n in NAMES:
            t = getattr(b, n)
            setattr(p, n, t.data_ptr() + off * t.element_size())
        return p

    def migrate(self, stream=None):
        b = self.batch
        s = ReplicaBatch._stream(stream)
        for r, spec in enumerate(b.specs):
            h = self.halos[r]
            off = b.offsets[r]
            _abi.check(_abi.lib().dsdx_migrate_wrap(C.byref(h._box), b.x.data_ptr() + 8 * off,
                                                    b.y.data_ptr() + 8 * off,
                                                    b.z.data_ptr() + 8 * off, spec.owned, s))
            h.reset()

    def exchange_halo(self, stream=None):
        b = self.batch
        s = ReplicaBatch._stream(stream)
        for r, spec in enumerate(b.specs):
            h = self.halos[r]
            p = self._view(r)
            g = C.c_uint64()
            _abi.check(_abi.lib().dsdx_halo_build(h.plan, C.byref(h._box), C.byref(p), None, 0,
                                                  spec.owned, b.capacities[r], s, C.byref(g)))
            self.ghosts[r] = g.value
            h.owned, h.ghosts = spec.owned, g.value

    def reverse_halo(self, stream=None):
        b = self.batch
        s = ReplicaBatch._stream(stream)
        for r, spec in enumerate(b.specs):
            off = 8 * b.offsets[r]
            _abi.check(_abi.lib().dsdx_reverse_halo(self.halos[r].plan, b.fx.data_ptr() + off,
                                                    b.fy.data_ptr() + off, b.fz.data_ptr() + off,
                                                    spec.owned, s))


class CapturedStep:
    """A step recorded by dsdx_graph_begin / dsdx_graph_end (a HIP graph the
    library instantiates); replay() launches it on the current stream."""

    def __init__(self, graph):
        self._graph = graph

    def replay(self, stream=None) -> None:
        _abi.check(_abi.lib().dsdx_graph_launch(self._graph, ReplicaBatch._stream(stream)),
                   "graph launch")

    def close(self) -> None:
        if self._graph:
            _abi.check(_abi.lib().dsdx_graph_destroy(self._graph), "graph destroy")
            self._graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def capture_step(batch: ReplicaBatch, app_before_halo=None, app_before_reverse=None,
                 warmup: int = 2, fused: bool = False) -> CapturedStep:
    """Captures migrate -> exchange_halo -> [app] -> reverse_halo of every
    replica in one HIP graph through the library (dsdx_graph_begin /
    dsdx_graph_end) and returns it.  The application callbacks
    (displacement, forces) are recorded in place; they run on the capturing
    stream and must not allocate.  `warmup` steps run eagerly first.  With
    `fused` migrate and the rebuild share one pass over the positions."""
    _abi.timing_enable(False)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())

    def body():
        if app_before_halo:
            app_before_halo()
        if fused:
            batch.migrate_exchange_halo()
        else:
            batch.migrate()
            batch.exchange_halo()
        if app_before_reverse:
            app_before_reverse()
        batch.reverse_halo()

    with torch.cuda.stream(s):
        for _ in range(warmup):
            body()
    s.synchronize()
    graph = C.c_void_p()
    with torch.cuda.stream(s):
        _abi.check(_abi.lib().dsdx_graph_begin(C.c_void_p(s.cuda_stream)), "graph begin")
        try:
            body()
        finally:
            rc = _abi.lib().dsdx_graph_end(C.c_void_p(s.cuda_stream), C.byref(graph))
        _abi.check(rc, "graph end")
    torch.cuda.current_stream().wait_stream(s)
    return CapturedStep(graph)
