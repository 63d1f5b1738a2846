"""Process grids on one GPU: the subdomains of a Cartesian decomposition of
one box, each handled as a replica of the halo kernels, with the semantics
of the reference's Host path on that many MPI ranks (see ``dsdx_grid_*`` in
include/dsdx_hip.h).

Storage: two sets of SoA device arrays (``DeviceParticles``) of
``ranks * capacity_per_rank`` slots each; subdomain r owns slots
[r * capacity_per_rank, (r + 1) * capacity_per_rank), owned particles first,
then its ghosts.  ``migrate`` writes the migrated particles into the spare
set and makes it current.  Fields registered with ``add_field`` (the
reference's ParticleInfo::add_field, include/dsdecompx.hpp:105) are held the same
way, travel with their particles in ``migrate`` and into the ghosts in
``exchange_halo``, and are exchanged by ``forward_exchange`` /
``reverse_exchange``.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .halo import BoxConfig, DeviceParticles, _field_array, field_struct


class SubdomainGrid:
    def __init__(self, box: BoxConfig, dims, capacity_per_rank: int, device="cuda", stream=None):
        self.box = box
        self.dims = tuple(int(d) for d in dims)
        self.ranks = self.dims[0] * self.dims[1] * self.dims[2]
        self.capacity = int(capacity_per_rank)
        self.device = torch.device(device)
        self._storage = [DeviceParticles(self.ranks * self.capacity, self.device) for _ in range(2)]
        self._current = 0
        self._fields = {}   # name -> ([tensor, tensor], op)
        self._box = box.c()
        self._grid = C.c_void_p()
        d = (C.c_int32 * 3)(*self.dims)
        _abi.check(_abi.lib().dsdx_grid_create(C.byref(self._box), d, self.capacity,
                                               self._stream(stream), C.byref(self._grid)),
                   "grid create")

    @staticmethod
    def _stream(stream):
        return (stream or torch.cuda.current_stream()).cuda_stream

    def close(self):
        if self._grid:
            _abi.lib().dsdx_grid_destroy(self._grid)
            self._grid = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def particles(self) -> DeviceParticles:
        """The current storage."""
        return self._storage[self._current]

    def add_field(self, name: str, dtype=torch.float64, components: int = 1, op: int = _abi.SUM):
        """Registers a per-particle field of `components` values (zeros)."""
        if name in self._fields:
            raise ValueError(f"field {name!r} is already registered")
        shape = (self.ranks * self.capacity,) if components == 1 else (self.ranks * self.capacity,
                                                                        components)
        self._fields[name] = ([torch.zeros(shape, dtype=dtype, device=self.device) for _ in range(2)],
                              int(op))

    def field(self, name: str) -> torch.Tensor:
        """The current storage of a registered field."""
        return self._fields[name][0][self._current]

    def _table(self, names=None, side=None):
        side = self._current if side is None else side
        names = list(self._fields) if names is None else list(names)
        return _field_array([field_struct(self._fields[n][0][side], self._fields[n][1]) for n in names])

    def slots(self, r: int, count: int) -> slice:
        return slice(r * self.capacity, r * self.capacity + count)

    def set_owned(self, owned, stream=None) -> None:
        arr = (C.c_uint64 * self.ranks)(*[int(v) for v in owned])
        _abi.check(_abi.lib().dsdx_grid_set_owned(self._grid, arr, self._stream(stream)),
                   "grid set owned")

    def counts(self):
        """(owned, ghosts, owners) of every subdomain: ghosts and distinct
        boundary owners of the last halo build."""
        o, g, u = (C.c_uint64 * self.ranks)(), (C.c_uint64 * self.ranks)(), (C.c_uint64 * self.ranks)()
        _abi.check(_abi.lib().dsdx_grid_counts(self._grid, o, g, u))
        return list(o), list(g), list(u)

    def migrate(self, stream=None) -> list:
        """Migration on every subdomain; returns the new owned counts."""
        src = self._storage[self._current].c()
        dst = self._storage[1 - self._current].c()
        ff, n = self._table(side=self._current)
        ft, _ = self._table(side=1 - self._current)
        _abi.check(_abi.lib().dsdx_grid_migrate(self._grid, C.byref(src), C.byref(dst), ff, ft, n,
                                                self._stream(stream)), "grid migrate")
        self._current = 1 - self._current
        return self.counts()[0]

    def exchange_halo(self, stream=None) -> list:
        """Halo rebuild on every subdomain; returns the ghost counts."""
        p = self.particles.c()
        g = (C.c_uint64 * self.ranks)()
        arr, n = self._table()
        _abi.check(_abi.lib().dsdx_grid_halo_build(self._grid, C.byref(p), arr, n,
                                                   self._stream(stream), g), "grid halo")
        return list(g)

    def refresh(self, positions_only: bool = False, stream=None) -> None:
        """Ghosts re-read their owners along the last build's routes."""
        p = self.particles.c()
        arr, n = self._table()
        _abi.check(_abi.lib().dsdx_grid_refresh(self._grid, C.byref(p), arr, n,
                                                1 if positions_only else 0, self._stream(stream)),
                   "grid refresh")

    def reverse_exchange(self, names=None, stream=None) -> None:
        """Ghost values of the registered fields reduced into their owners."""
        arr, n = self._table(names)
        _abi.check(_abi.lib().dsdx_grid_reverse_exchange(self._grid, arr, n, self._stream(stream)),
                   "grid reverse exchange")

    def forward_exchange(self, names=None, stream=None) -> None:
        """Owner values of the registered fields copied into their ghosts."""
        arr, n = self._table(names)
        _abi.check(_abi.lib().dsdx_grid_forward_exchange(self._grid, arr, n, self._stream(stream)),
                   "grid forward exchange")

    def reverse_halo(self, stream=None) -> None:
        p = self.particles
        _abi.check(_abi.lib().dsdx_grid_reverse_halo(self._grid, p.fx.data_ptr(), p.fy.data_ptr(),
                                                     p.fz.data_ptr(), self._stream(stream)),
                   "grid reverse halo")
