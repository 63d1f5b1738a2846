"""ctypes front end of the C oracle (oracle/dsdx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker and by ``bench.py``'s cpu_baseline leg.  The product package
``dsdecompx_amd`` never imports it.

All arrays are numpy arrays on the host.  ``Particles`` mirrors the SoA layout
of the reference's ``ParticleView`` (include/field.hpp:113-131): 11 float64
arrays and 2 int32 arrays, each of length ``capacity`` >= owned + ghosts.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_dsdx.so")

MAX_NEIGHBORS = 26

# FieldType / ReduceOp of include/field.hpp:22-36
INT32, INT64, UINT32, UINT64, FLOAT32, FLOAT64 = range(6)
SUM, MIN, MAX, REPLACE = range(4)
NUMPY_TYPE = {INT32: np.int32, INT64: np.int64, UINT32: np.uint32,
              UINT64: np.uint64, FLOAT32: np.float32, FLOAT64: np.float64}
FIELD_TYPE = {np.dtype(v): k for k, v in NUMPY_TYPE.items()}


class _Box(C.Structure):
    _fields_ = [("lower", C.c_double * 3), ("upper", C.c_double * 3),
                ("periodic", C.c_int32 * 3), ("width", C.c_double),
                ("edges_and_corners", C.c_int32),
                ("local_lower", C.c_double * 3), ("local_upper", C.c_double * 3),
                ("local_given", C.c_int32)]


class _Neighbors(C.Structure):
    _fields_ = [("count", C.c_int32),
                ("direction", (C.c_int32 * 3) * MAX_NEIGHBORS),
                ("shift", (C.c_double * 3) * MAX_NEIGHBORS),
                ("slot", C.c_int32 * MAX_NEIGHBORS),
                ("active", C.c_int32 * MAX_NEIGHBORS)]


class _Plan(C.Structure):
    _fields_ = [("send_count", C.c_int64 * MAX_NEIGHBORS),
                ("source", C.POINTER(C.c_int32) * MAX_NEIGHBORS),
                ("send_capacity", C.c_int64 * MAX_NEIGHBORS),
                ("recv_count", C.c_int64 * MAX_NEIGHBORS),
                ("recv_from", C.c_int32 * MAX_NEIGHBORS),
                ("ghost_total", C.c_int64)]


MAX_RANKS = 64


class _Grid(C.Structure):
    _fields_ = [("global_", _Box), ("dims", C.c_int32 * 3), ("ranks", C.c_int32),
                ("cuts", (C.c_double * (MAX_RANKS + 1)) * 3),
                ("coord", (C.c_int32 * 3) * MAX_RANKS),
                ("local", _Box * MAX_RANKS),
                ("nb", _Neighbors * MAX_RANKS),
                ("nb_rank", (C.c_int32 * MAX_NEIGHBORS) * MAX_RANKS)]


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


class _Particles(C.Structure):
    _fields_ = [(n, _DP) for n in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy",
                                   "fz", "mass", "charge")] + \
               [("type", _IP), ("global_id", _IP)]


class _Field(C.Structure):
    _fields_ = [("data", C.c_void_p), ("type", C.c_int32),
                ("components", C.c_int32), ("op", C.c_int32)]


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.dsdxo_direction_slot.restype = C.c_int
        L.dsdxo_opposite_slot.restype = C.c_int
        L.dsdxo_neighbors_build.argtypes = [C.POINTER(_Box), C.POINTER(_Neighbors)]
        L.dsdxo_halo_build.argtypes = [C.POINTER(_Box), C.POINTER(_Neighbors), C.c_int64,
                                       _DP, _DP, _DP, C.POINTER(_Plan)]
        L.dsdxo_halo_build.restype = C.c_int
        L.dsdxo_plan_free.argtypes = [C.POINTER(_Plan)]
        L.dsdxo_ghost_owners.argtypes = [C.POINTER(_Neighbors), C.POINTER(_Plan), _IP]
        L.dsdxo_halo_write_ghosts.argtypes = [C.POINTER(_Neighbors), C.POINTER(_Plan),
                                              C.c_int64, C.POINTER(_Particles),
                                              C.POINTER(_Field), C.c_int32]
        L.dsdxo_reverse_exchange.argtypes = [C.POINTER(_Neighbors), C.POINTER(_Plan),
                                             C.c_int64, C.POINTER(_Field), C.c_int32]
        L.dsdxo_reverse_exchange.restype = C.c_int
        L.dsdxo_reverse_halo.argtypes = [C.POINTER(_Neighbors), C.POINTER(_Plan),
                                         C.c_int64, _DP, _DP, _DP]
        L.dsdxo_reverse_halo.restype = C.c_int
        L.dsdxo_forward_exchange.argtypes = [C.POINTER(_Neighbors), C.POINTER(_Plan),
                                             C.c_int64, C.POINTER(_Field), C.c_int32]
        L.dsdxo_migrate_wrap.argtypes = [C.POINTER(_Box), C.c_int64, _DP, _DP, _DP]
        L.dsdxo_local_domain.argtypes = [C.POINTER(_Box), _DP, _DP]
        L.dsdxo_rescale_cuts.argtypes = [_DP, _DP, _DP, _DP, _DP, _DP]
        L.dsdxo_grid_build.argtypes = [C.POINTER(_Box), C.POINTER(C.c_int32), C.c_int32,
                                       C.POINTER(_Grid)]
        L.dsdxo_grid_build.restype = C.c_int
        L.dsdxo_grid_owner_rank.argtypes = [C.POINTER(_Grid), C.c_double, C.c_double, C.c_double]
        L.dsdxo_grid_owner_rank.restype = C.c_int32
        L.dsdxo_grid_migrate.argtypes = [C.POINTER(_Grid), C.POINTER(_Particles),
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                         C.POINTER(_Field), C.c_int32,
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.dsdxo_grid_migrate.restype = C.c_int
        L.dsdxo_grid_halo.argtypes = [C.POINTER(_Grid), C.POINTER(_Particles),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(_Plan), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int64), C.POINTER(_Field), C.c_int32]
        L.dsdxo_grid_halo.restype = C.c_int
        L.dsdxo_grid_reverse_exchange.argtypes = [C.POINTER(_Grid), C.POINTER(_Plan),
                                                  C.POINTER(C.c_int64), C.POINTER(_Field),
                                                  C.c_int32]
        L.dsdxo_grid_reverse_exchange.restype = C.c_int
        L.dsdxo_grid_forward_exchange.argtypes = [C.POINTER(_Grid), C.POINTER(_Plan),
                                                  C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                                  C.POINTER(_Field), C.c_int32]
        L.dsdxo_grid_forward_exchange.restype = C.c_int
        _U8 = C.POINTER(C.c_ubyte)
        N, PL = C.POINTER(_Neighbors), C.POINTER(_Plan)
        L.dsdxo_wire_layout.argtypes = [C.POINTER(_Field), C.c_int32, C.c_int32,
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.dsdxo_halo_pack.argtypes = [N, PL, C.POINTER(_Particles), C.POINTER(_Field), C.c_int32,
                                      C.c_int32, _U8]
        L.dsdxo_halo_pack.restype = C.c_int64
        L.dsdxo_unpack_records.argtypes = [_U8, C.c_int64, C.POINTER(_Particles), C.c_int64,
                                           C.POINTER(_Field), C.c_int32]
        L.dsdxo_unpack_positions.argtypes = [_DP, C.c_int64, C.POINTER(_Particles), C.c_int64]
        L.dsdxo_pack_reverse_records.argtypes = [_DP, _DP, _DP, C.c_int64, C.c_int64, _U8]
        L.dsdxo_accumulate_reverse.argtypes = [N, PL, _U8, _DP, _DP, _DP]
        L.dsdxo_pack_field_records.argtypes = [C.POINTER(_Field), C.c_int32, C.c_int64,
                                               C.c_int64, _U8]
        L.dsdxo_scatter_field_records.argtypes = [C.POINTER(_Field), C.c_int32, _U8, C.c_int64,
                                                  C.c_int64]
        L.dsdxo_gather_field_records.argtypes = [N, PL, C.POINTER(_Field), C.c_int32, _U8]
        L.dsdxo_gather_field_records.restype = C.c_int64
        L.dsdxo_reduce_field_records.argtypes = [N, PL, C.POINTER(_Field), C.c_int32, _U8]
        L.dsdxo_loopback_exchange.argtypes = [N, PL, _U8, _U8, C.c_int64, C.c_int32]
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def _ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)


@dataclass
class Box:
    """Global domain + halo configuration (include/domain.hpp:11-30).
    `local` = (lower, upper) of the local domain whose faces the slab tests
    use; None: the cuts of a fresh 1x1x1 topology, lower + extent * i / 1
    (src/topology.cpp:66-72)."""
    lower: tuple = (0.0, 0.0, 0.0)
    upper: tuple = (1.0, 1.0, 1.0)
    periodic: tuple = (True, True, True)
    width: float = 0.0
    edges_and_corners: bool = True
    local: tuple = None

    def c(self) -> _Box:
        b = _Box()
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


def local_domain(box: "Box"):
    """The local domain the slab tests of `box` use (dsdxo_local_domain)."""
    lo, hi = np.zeros(3), np.zeros(3)
    lib().dsdxo_local_domain(C.byref(box.c()), _dp(lo), _dp(hi))
    return tuple(lo), tuple(hi)


def rescale_cuts(old: "Box", new: "Box", cuts):
    """rescale_domain's in-place cut update (dsdxo_rescale_cuts): `cuts` =
    (lower cuts, upper cuts) of the old domain -> those of the new one."""
    arrs = [np.array(v, np.float64) for v in (old.lower, old.upper, new.lower, new.upper)]
    lo, hi = np.array(cuts[0], np.float64), np.array(cuts[1], np.float64)
    lib().dsdxo_rescale_cuts(*[_dp(a) for a in arrs], _dp(lo), _dp(hi))
    return tuple(lo), tuple(hi)


@dataclass
class Neighbors:
    direction: np.ndarray
    shift: np.ndarray
    slot: np.ndarray
    active: np.ndarray
    raw: _Neighbors = field(repr=False)


def neighbors(box: Box) -> Neighbors:
    nb = _Neighbors()
    lib().dsdxo_neighbors_build(C.byref(box.c()), C.byref(nb))
    n = nb.count
    return Neighbors(
        direction=np.array([[nb.direction[i][a] for a in range(3)] for i in range(n)],
                           dtype=np.int32).reshape(n, 3),
        shift=np.array([[nb.shift[i][a] for a in range(3)] for i in range(n)],
                       dtype=np.float64).reshape(n, 3),
        slot=np.array([nb.slot[i] for i in range(n)], dtype=np.int32),
        active=np.array([nb.active[i] for i in range(n)], dtype=bool),
        raw=nb)


class Particles:
    """Host SoA storage of capacity ``capacity`` (owned first, then ghosts)."""
    F64 = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "charge")
    I32 = ("type", "global_id")

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        for n in self.F64:
            setattr(self, n, np.zeros(self.capacity, np.float64))
        for n in self.I32:
            setattr(self, n, np.zeros(self.capacity, np.int32))

    def grow(self, capacity: int, keep: int) -> None:
        if capacity <= self.capacity:
            return
        for n in self.F64 + self.I32:
            old = getattr(self, n)
            new = np.zeros(capacity, old.dtype)
            new[:keep] = old[:keep]
            setattr(self, n, new)
        self.capacity = capacity

    def c(self, omit=()) -> _Particles:
        p = _Particles()
        for n in self.F64:
            setattr(p, n, None if n in omit else _dp(getattr(self, n)))
        for n in self.I32:
            setattr(p, n, None if n in omit else _ip(getattr(self, n)))
        return p


class Field:
    """A registered application field: AoS array [capacity, components]."""

    def __init__(self, data: np.ndarray, op: int = SUM):
        assert data.flags.c_contiguous
        self.data = data
        self.type = FIELD_TYPE[data.dtype]
        self.components = 1 if data.ndim == 1 else int(np.prod(data.shape[1:]))
        self.op = op

    def c(self) -> _Field:
        return _Field(self.data.ctypes.data, self.type, self.components, self.op)


def _fields_c(fields):
    fields = list(fields or [])
    arr = (_Field * max(1, len(fields)))()
    for i, f in enumerate(fields):
        arr[i] = f.c()
    return arr, len(fields)


class Plan:
    """The Host halo plan of the last build (src/halo.cpp:164-238)."""

    def __init__(self):
        self.raw = _Plan()

    def __del__(self):
        try:
            lib().dsdxo_plan_free(C.byref(self.raw))
        except Exception:
            pass

    @property
    def ghost_total(self) -> int:
        return int(self.raw.ghost_total)

    def send_counts(self, nb: Neighbors) -> np.ndarray:
        return np.array([self.raw.send_count[n] for n in range(len(nb.slot))], np.int64)

    def recv_counts(self, nb: Neighbors) -> np.ndarray:
        return np.array([self.raw.recv_count[n] for n in range(len(nb.slot))], np.int64)

    def send_list(self, n: int) -> np.ndarray:
        c = int(self.raw.send_count[n])
        if c == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self.raw.source[n], shape=(c,)).copy()


class Oracle:
    """Host-path restatement bound to one box (a Decomp on a 1x1x1 grid)."""

    def __init__(self, box: Box):
        self.box = box
        self._box = box.c()
        self.nb = neighbors(box)
        self.plan = Plan()

    def migrate(self, p: Particles, owned: int) -> None:
        lib().dsdxo_migrate_wrap(C.byref(self._box), owned, _dp(p.x), _dp(p.y), _dp(p.z))

    def build(self, p: Particles, owned: int) -> int:
        rc = lib().dsdxo_halo_build(C.byref(self._box), C.byref(self.nb.raw), owned,
                                     _dp(p.x), _dp(p.y), _dp(p.z), C.byref(self.plan.raw))
        if rc != 0:
            raise MemoryError("oracle halo build failed")
        return self.plan.ghost_total

    def write_ghosts(self, p: Particles, owned: int, fields=(), omit=()) -> None:
        arr, n = _fields_c(fields)
        lib().dsdxo_halo_write_ghosts(C.byref(self.nb.raw), C.byref(self.plan.raw), owned,
                                      C.byref(p.c(omit)), arr, n)

    def exchange_halo(self, p: Particles, owned: int, fields=(), omit=()) -> int:
        g = self.build(p, owned)
        if owned + g > p.capacity:
            raise ValueError("particle storage too small for the ghosts")
        self.write_ghosts(p, owned, fields, omit)
        return g

    def ghost_owners(self) -> np.ndarray:
        out = np.zeros(max(1, self.plan.ghost_total), np.int32)
        lib().dsdxo_ghost_owners(C.byref(self.nb.raw), C.byref(self.plan.raw), _ip(out))
        return out[:self.plan.ghost_total]

    def reverse_halo(self, p: Particles, owned: int) -> None:
        rc = lib().dsdxo_reverse_halo(C.byref(self.nb.raw), C.byref(self.plan.raw), owned,
                                      _dp(p.fx), _dp(p.fy), _dp(p.fz))
        if rc != 0:
            raise RuntimeError("oracle reverse halo failed")

    def reverse_exchange(self, owned: int, fields) -> None:
        arr, n = _fields_c(fields)
        if lib().dsdxo_reverse_exchange(C.byref(self.nb.raw), C.byref(self.plan.raw),
                                        owned, arr, n) != 0:
            raise RuntimeError("oracle reverse exchange failed")

    def forward_exchange(self, owned: int, fields) -> None:
        arr, n = _fields_c(fields)
        lib().dsdxo_forward_exchange(C.byref(self.nb.raw), C.byref(self.plan.raw),
                                     owned, arr, n)


# --- wire format (dsdx_oracle_wire.c) -------------------------------------------
WIRE_PARTICLE, WIRE_EXCHANGE = 0, 1
RECORD_BYTES = 96


def _u8(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_ubyte))


def wire_layout(fields, kind: int):
    """(field byte offsets, record stride) of a particle record (kind 0) or an
    exchange record (kind 1)."""
    arr, n = _fields_c(fields)
    offs = (C.c_int64 * max(1, n))()
    stride = C.c_int64()
    lib().dsdxo_wire_layout(arr, n, kind, offs, C.byref(stride))
    return [int(offs[f]) for f in range(n)], int(stride.value)


class Wire:
    """The device path's wire steps on an ``Oracle``'s last plan: buffers are
    numpy uint8 arrays; record counts are the plan's send / receive totals."""

    def __init__(self, ora: "Oracle"):
        self.ora = ora

    @property
    def sent(self) -> int:
        return int(sum(self.ora.plan.raw.send_count[n] for n in range(self.ora.nb.raw.count)))

    def pack(self, p: Particles, fields=(), positions_only=False) -> np.ndarray:
        stride = 24 if positions_only else wire_layout(fields, WIRE_PARTICLE)[1]
        out = np.zeros(max(1, self.sent) * stride, np.uint8)
        arr, n = _fields_c(fields)
        lib().dsdxo_halo_pack(C.byref(self.ora.nb.raw), C.byref(self.ora.plan.raw), C.byref(p.c()),
                              arr, n, 1 if positions_only else 0, _u8(out))
        return out[:self.sent * stride]

    def exchange(self, send: np.ndarray, record_bytes: int, reverse=False) -> np.ndarray:
        recv = np.zeros(max(1, len(send)), np.uint8)
        src = np.ascontiguousarray(send) if len(send) else np.zeros(1, np.uint8)
        lib().dsdxo_loopback_exchange(C.byref(self.ora.nb.raw), C.byref(self.ora.plan.raw),
                                      _u8(src), _u8(recv), record_bytes, 1 if reverse else 0)
        return recv[:len(send)]

    @staticmethod
    def unpack(records: np.ndarray, count: int, p: Particles, dest: int, fields=()) -> None:
        arr, n = _fields_c(fields)
        rec = records if len(records) else np.zeros(1, np.uint8)
        lib().dsdxo_unpack_records(_u8(rec), count, C.byref(p.c()), dest, arr, n)

    @staticmethod
    def unpack_positions(records: np.ndarray, count: int, p: Particles, dest: int) -> None:
        pos = np.frombuffer(records.tobytes() or bytes(8), np.float64).copy()
        lib().dsdxo_unpack_positions(_dp(pos), count, C.byref(p.c()), dest)

    @staticmethod
    def pack_reverse(p: Particles, first: int, count: int) -> np.ndarray:
        out = np.zeros(max(1, count) * RECORD_BYTES, np.uint8)
        lib().dsdxo_pack_reverse_records(_dp(p.fx), _dp(p.fy), _dp(p.fz), first, count, _u8(out))
        return out[:count * RECORD_BYTES]

    def accumulate_reverse(self, records: np.ndarray, p: Particles) -> None:
        rec = records if len(records) else np.zeros(1, np.uint8)
        lib().dsdxo_accumulate_reverse(C.byref(self.ora.nb.raw), C.byref(self.ora.plan.raw),
                                       _u8(rec), _dp(p.fx), _dp(p.fy), _dp(p.fz))

    @staticmethod
    def pack_fields(fields, first: int, count: int) -> np.ndarray:
        stride = wire_layout(fields, WIRE_EXCHANGE)[1]
        out = np.zeros(max(1, count) * stride, np.uint8)
        arr, n = _fields_c(fields)
        lib().dsdxo_pack_field_records(arr, n, first, count, _u8(out))
        return out[:count * stride]

    @staticmethod
    def scatter_fields(fields, records: np.ndarray, count: int, first: int) -> None:
        arr, n = _fields_c(fields)
        rec = records if len(records) else np.zeros(1, np.uint8)
        lib().dsdxo_scatter_field_records(arr, n, _u8(rec), count, first)

    def gather_fields(self, fields) -> np.ndarray:
        stride = wire_layout(fields, WIRE_EXCHANGE)[1]
        out = np.zeros(max(1, self.sent) * stride, np.uint8)
        arr, n = _fields_c(fields)
        lib().dsdxo_gather_field_records(C.byref(self.ora.nb.raw), C.byref(self.ora.plan.raw),
                                         arr, n, _u8(out))
        return out[:self.sent * stride]

    def reduce_fields(self, fields, records: np.ndarray) -> None:
        arr, n = _fields_c(fields)
        rec = records if len(records) else np.zeros(1, np.uint8)
        lib().dsdxo_reduce_field_records(C.byref(self.ora.nb.raw), C.byref(self.ora.plan.raw),
                                         arr, n, _u8(rec))


class Grid:
    """Host path on a dims[0] x dims[1] x dims[2] process grid, every rank
    held here (src/topology.cpp, src/halo.cpp, src/migration.cpp at P ranks).
    ``lift_unsplit`` lifts the unsplit-axis skip (src/topology.cpp:118-127),
    which turns 1x1x1 into the self-loopback of ``Oracle``.

    ``parts[r]`` is rank r's ``Particles`` (owned first, then ghosts) and
    ``owned[r]`` its owned count; ``fields[r]`` its registered fields (the
    same schema on every rank)."""

    def __init__(self, box: Box, dims, lift_unsplit: bool = False):
        self.box = box
        self.raw = _Grid()
        d = (C.c_int32 * 3)(*[int(v) for v in dims])
        if lib().dsdxo_grid_build(C.byref(box.c()), d, 1 if lift_unsplit else 0,
                                  C.byref(self.raw)) != 0:
            raise ValueError("bad process grid")
        self.ranks = int(self.raw.ranks)
        self.dims = tuple(int(v) for v in dims)
        self.plans = (_Plan * self.ranks)()
        self.recv_rank = np.full(self.ranks * MAX_NEIGHBORS, -1, np.int32)
        self.ghosts = np.zeros(self.ranks, np.int64)

    def __del__(self):
        try:
            for r in range(self.ranks):
                lib().dsdxo_plan_free(C.byref(self.plans[r]))
        except Exception:
            pass

    # --- topology -------------------------------------------------------------
    def local(self, r: int):
        b = self.raw.local[r]
        return tuple(b.lower), tuple(b.upper)

    def coord(self, r: int):
        return tuple(self.raw.coord[r])

    def neighbor_list(self, r: int):
        """[(direction, slot, rank, shift)] in list order."""
        nb = self.raw.nb[r]
        return [(tuple(nb.direction[n]), int(nb.slot[n]), int(self.raw.nb_rank[r][n]),
                 tuple(nb.shift[n])) for n in range(nb.count)]

    def cuts(self, a: int):
        return [self.raw.cuts[a][i] for i in range(self.dims[a] + 1)]

    def owner_rank(self, x, y, z) -> int:
        return int(lib().dsdxo_grid_owner_rank(C.byref(self.raw), x, y, z))

    # --- operations -------------------------------------------------------------
    def _parts(self, parts):
        arr = (_Particles * self.ranks)()
        for r, p in enumerate(parts):
            arr[r] = p.c()
        return arr

    def _fields(self, fields):
        nf = len(fields[0]) if fields else 0
        arr = (_Field * max(1, self.ranks * nf))()
        for r in range(self.ranks if nf else 0):
            for f in range(nf):
                arr[r * nf + f] = fields[r][f].c()
        return arr, nf

    def migrate(self, parts, owned, fields=None):
        cap = (C.c_int64 * self.ranks)(*[p.capacity for p in parts])
        own = (C.c_int64 * self.ranks)(*[int(v) for v in owned])
        sent = (C.c_int64 * self.ranks)()
        recv = (C.c_int64 * self.ranks)()
        arr, nf = self._fields(fields)
        if lib().dsdxo_grid_migrate(C.byref(self.raw), self._parts(parts), cap, own, arr, nf,
                                    sent, recv) != 0:
            raise ValueError("oracle migration failed (capacity)")
        for r in range(self.ranks):
            owned[r] = int(own[r])
        return list(sent), list(recv)

    def exchange_halo(self, parts, owned, fields=None):
        cap = (C.c_int64 * self.ranks)(*[p.capacity for p in parts])
        own = (C.c_int64 * self.ranks)(*[int(v) for v in owned])
        g = (C.c_int64 * self.ranks)()
        arr, nf = self._fields(fields)
        rc = lib().dsdxo_grid_halo(C.byref(self.raw), self._parts(parts), cap, own, self.plans,
                                   self.recv_rank.ctypes.data_as(C.POINTER(C.c_int32)), g, arr, nf)
        self.ghosts[:] = list(g)
        if rc != 0:
            raise ValueError("oracle halo failed (capacity)")
        return [int(v) for v in g]

    def send_counts(self, r: int):
        nb = self.raw.nb[r]
        return [int(self.plans[r].send_count[n]) for n in range(nb.count)]

    def reverse_exchange(self, owned, fields):
        own = (C.c_int64 * self.ranks)(*[int(v) for v in owned])
        arr, nf = self._fields(fields)
        if lib().dsdxo_grid_reverse_exchange(C.byref(self.raw), self.plans, own, arr, nf) != 0:
            raise RuntimeError("oracle reverse exchange failed")

    def reverse_halo(self, parts, owned):
        fields = [[Field(p.fx), Field(p.fy), Field(p.fz)] for p in parts]
        self.reverse_exchange(owned, fields)

    def forward_exchange(self, owned, fields):
        own = (C.c_int64 * self.ranks)(*[int(v) for v in owned])
        arr, nf = self._fields(fields)
        lib().dsdxo_grid_forward_exchange(C.byref(self.raw), self.plans,
                                          self.recv_rank.ctypes.data_as(C.POINTER(C.c_int32)),
                                          own, arr, nf)
