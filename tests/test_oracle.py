"""Pins the CPU oracle against what fixes the reference's Host semantics
independently of the oracle's own code: closed forms, hand-derived known
answers, the reference examples' expected outputs, exact-rational brute force
and the invariants the reference's own tests assert
(examples/sanity_check/mpi/regression.cpp:46-101,
examples/sanity_check/cuda/forward_exchange.cu:103-167)."""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle as po


def _slot(d):
    raw = (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1)
    return raw if raw < 13 else raw - 1


def _make(box, pos, extra=0):
    n = len(pos)
    p = po.Particles(n * 27 + extra + 1)
    pos = np.asarray(pos, np.float64).reshape(n, 3)
    p.x[:n], p.y[:n], p.z[:n] = pos[:, 0], pos[:, 1], pos[:, 2]
    p.global_id[:n] = np.arange(n, dtype=np.int32)
    p.mass[:n] = 1.0
    return p


# ---------------------------------------------------------------- topology
def test_neighbor_table_closed_form(oracle_lib):
    L = np.array([7.5, 9.0, 11.25])
    nb = po.neighbors(po.Box(upper=tuple(L), width=1.0))
    dirs = [d for d in itertools.product((-1, 0, 1), repeat=3) if d != (0, 0, 0)]
    assert len(nb.slot) == 26
    assert [tuple(d) for d in nb.direction] == dirs       # dx, dy, dz nested order
    assert list(nb.slot) == list(range(26))
    assert nb.active.all()
    for k, d in enumerate(dirs):
        # periodic image of a send toward d is shifted by -d * extent
        assert tuple(nb.shift[k]) == tuple(-np.array(d) * L + 0.0)
        opp = tuple(-x for x in d)
        assert _slot(opp) == 25 - k
        assert oracle_lib.lib().dsdxo_opposite_slot(*d) == 25 - k


def test_neighbor_table_nonperiodic_and_faces(oracle_lib):
    nb = po.neighbors(po.Box(upper=(4.0, 4.0, 4.0), periodic=(True, True, False), width=0.5))
    active = [tuple(d) for d, a in zip(nb.direction, nb.active) if a]
    assert len(active) == 8 and all(d[2] == 0 for d in active)
    faces = po.neighbors(po.Box(upper=(4.0, 4.0, 4.0), width=0.5, edges_and_corners=False))
    assert list(faces.slot) == [4, 10, 12, 13, 15, 21]


# ---------------------------------------------------------------- halo build
def test_reference_example_corner_particle(oracle_lib):
    """examples/getting_started/main.cpp: particle at lower+0.05 on every axis,
    box 8, width 0.25.  Every rank of a P-rank run receives 2^k-1 ghosts for k
    split axes (examples/README.md:10-16: 1 on 2 ranks); on the loopback grid
    all three axes wrap, so 7 ghosts return fx = 7."""
    box = po.Box(upper=(8.0, 8.0, 8.0), width=0.25)
    o = po.Oracle(box)
    p = _make(box, [(0.05, 0.05, 0.05)])
    o.migrate(p, 1)
    g = o.exchange_halo(p, 1)
    assert g == 7
    p.fx[:] = 0.0
    p.fx[1:1 + g] = 1.0
    o.reverse_halo(p, 1)
    assert p.fx[0] == 7.0
    # registered_field example: reverse_exchange(Sum) of 1.0 per ghost equals it
    contribution = np.zeros(p.capacity)
    o.exchange_halo(p, 1)
    contribution[1:1 + g] = 1.0
    o.reverse_exchange(1, [po.Field(contribution, po.SUM)])
    assert contribution[0] == p.fx[0]


def test_exact_ties_follow_host_comparisons(oracle_lib):
    """Host uses px < lower+w and px > upper-w (src/halo.cpp:200-205)."""
    L, w = 10.0, 0.3
    lo_t, hi_t = 0.0 + w, L - w
    box = po.Box(upper=(L, L, L), width=w)
    cases = [((lo_t, 5.0, 5.0), 0),
             ((np.nextafter(lo_t, -1.0), 5.0, 5.0), 1),
             ((hi_t, 5.0, 5.0), 0),
             ((np.nextafter(hi_t, 20.0), 5.0, 5.0), 1),
             ((5.0, 5.0, 0.0), 1),
             ((5.0, 5.0, L), 1)]
    for pos, expected in cases:
        o = po.Oracle(box)
        assert o.exchange_halo(_make(box, [pos]), 1) == expected, pos


def _rn(q):
    """The double nearest to the rational q (ties to even): int / int true
    division in CPython is correctly rounded."""
    q = Fraction(q)
    return q.numerator / q.denominator


def test_upper_cut_is_lower_plus_extent(oracle_lib):
    """The slab tests use the local domain's faces, the cuts lower + extent *
    i / 1 (src/topology.cpp:66-72), not the domain's upper bound: for lower
    9.8, upper 53.6 the upper cut is one ulp below 53.6, so with w = 1.1 the
    high threshold is 52.49999999999999 and a particle at exactly 52.5 is in
    the high slab.  Each value below is derived with one exact rounding per
    operation (Fractions), not with the oracle's formula."""
    lo, up, w = 9.8, 53.6, 1.1
    extent = _rn(Fraction(up) - Fraction(lo))
    cut_hi = _rn(Fraction(lo) + Fraction(_rn(Fraction(extent) * 1) / 1))
    assert cut_hi == np.nextafter(up, 0.0)           # the cut differs from `upper`
    hi_t = _rn(Fraction(cut_hi) - Fraction(w))
    assert hi_t < 52.5 == _rn(Fraction(up) - Fraction(w))
    box = po.Box(lower=(lo,) * 3, upper=(up,) * 3, width=w)
    mid = 30.0
    for pos, expected in (((52.5, mid, mid), 1), ((hi_t, mid, mid), 0),
                          ((mid, 52.5, mid), 1), ((mid, mid, 52.5), 1),
                          ((_rn(Fraction(lo) + Fraction(w)), mid, mid), 0)):
        assert po.Oracle(box).exchange_halo(_make(box, [pos]), 1) == expected, pos
    # given cuts replace the derived ones: with the domain's own bounds 52.5 is
    # exactly on the threshold and has no image
    given = po.Box(lower=(lo,) * 3, upper=(up,) * 3, width=w, local=((lo,) * 3, (up,) * 3))
    assert po.Oracle(given).exchange_halo(_make(given, [(52.5, mid, mid)]), 1) == 0


def test_rescale_cuts_keep_their_fraction(oracle_lib):
    """rescale_domain moves every cut to new_lower + ((cut - old_lower) /
    old_extent) * new_extent (src/decomp.cpp:117-124), one rounding per
    operation.  Expected values from exact rationals; closed form: doubling a
    box at lower 0 doubles every cut.  The inputs separate this order from
    scaling by the extents' ratio, new_lower + (cut - old_lower) *
    (new_extent / old_extent)."""
    rng = np.random.default_rng(41)
    alt_differs = False
    for trial in range(200):
        lo_o = tuple(np.round(rng.uniform(-20, 20, 3), 2))
        up_o = tuple(np.round(np.array(lo_o) + rng.uniform(5, 90, 3), 2))
        lo_n = tuple(np.round(np.array(lo_o) + rng.uniform(-3, 3, 3), 2))
        up_n = tuple(np.round(np.array(lo_n) + rng.uniform(5, 90, 3), 2))
        old, new = po.Box(lower=lo_o, upper=up_o), po.Box(lower=lo_n, upper=up_n)
        cuts = po.local_domain(old)
        got = po.rescale_cuts(old, new, cuts)
        for a in range(3):
            oe = _rn(Fraction(up_o[a]) - Fraction(lo_o[a]))
            ne = _rn(Fraction(up_n[a]) - Fraction(lo_n[a]))
            assert cuts[0][a] == _rn(Fraction(lo_o[a]) + Fraction(_rn(Fraction(oe) * 0) / 1))
            assert cuts[1][a] == _rn(Fraction(lo_o[a]) + Fraction(_rn(Fraction(oe) * 1) / 1))
            for c in range(2):
                d = _rn(Fraction(cuts[c][a]) - Fraction(lo_o[a]))
                frac = _rn(Fraction(d) / Fraction(oe))
                want = _rn(Fraction(lo_n[a]) + Fraction(_rn(Fraction(frac) * Fraction(ne))))
                assert got[c][a] == want, (trial, a, c)
                ratio = _rn(Fraction(ne) / Fraction(oe))
                alt = _rn(Fraction(lo_n[a]) + Fraction(_rn(Fraction(d) * Fraction(ratio))))
                alt_differs |= alt != want
    assert alt_differs
    o2, n2 = po.Box(upper=(7.0, 9.5, 11.25)), po.Box(upper=(14.0, 19.0, 22.5))
    assert po.rescale_cuts(o2, n2, po.local_domain(o2)) == ((0.0,) * 3, (14.0, 19.0, 22.5))


def test_wide_halo_particle_in_both_slabs(oracle_lib):
    """w > L/2: a particle can be in the low and the high slab of an axis and
    then has images on both sides (26 for a particle in all six slabs)."""
    box = po.Box(upper=(2.0, 2.0, 2.0), width=1.5)
    o = po.Oracle(box)
    assert o.exchange_halo(_make(box, [(1.0, 1.0, 1.0)]), 1) == 26


def test_hand_derived_ghost_order(oracle_lib):
    """Receive block m holds send list opp(m) = 25 - m in owned order;
    blocks follow in slot order (src/runtime.cpp:946-957,
    src/backends/mpi_backend.cpp:366-375).  Expected list derived by hand."""
    L = 10.0
    box = po.Box(upper=(L, L, L), width=1.0)
    pos = [(5, 5, 5), (0.5, 5, 5), (9.5, 5, 5), (5, 0.5, 9.5), (0.2, 0.3, 0.4), (0.7, 5, 5)]
    p = _make(box, pos)
    o = po.Oracle(box)
    g = o.exchange_halo(p, 6)
    # (owner, shift) in ghost order
    expected = [(2, (-L, 0, 0)),                       # m=4  <- send slot 21 (+x)
                (3, (0, 0, -L)),                       # m=12 <- slot 13 (+z)
                (4, (0, 0, L)),                        # m=13 <- slot 12 (-z)
                (3, (0, L, -L)),                       # m=14 <- slot 11 (-y+z)
                (3, (0, L, 0)), (4, (0, L, 0)),        # m=15 <- slot 10 (-y)
                (4, (0, L, L)),                        # m=16 <- slot 9 (-y-z)
                (1, (L, 0, 0)), (4, (L, 0, 0)), (5, (L, 0, 0)),  # m=21 <- slot 4 (-x)
                (4, (L, 0, L)),                        # m=22 <- slot 3
                (4, (L, L, 0)),                        # m=24 <- slot 1
                (4, (L, L, L))]                        # m=25 <- slot 0
    assert g == len(expected)
    P = np.array(pos, np.float64)
    for k, (owner, s) in enumerate(expected):
        j = 6 + k
        assert p.global_id[j] == owner
        assert (p.x[j], p.y[j], p.z[j]) == tuple(P[owner] + np.array(s, np.float64))
        assert p.mass[j] == 1.0 and p.fx[j] == 0.0
    assert list(o.ghost_owners()) == [e[0] for e in expected]


def test_exact_rational_ghost_set(oracle_lib):
    """Geometric restatement in exact rationals: owner i has an image with
    shift -d*L iff for every axis with d != 0 the image lies inside the box
    grown by w on that side.  Independent of the oracle's slab/slot code."""
    rng = np.random.default_rng(7)
    for L, w, per in [(8.0, 1.3, (True,) * 3), (5.5, 0.9, (True, False, True)),
                      (3.0, 1.6, (True,) * 3)]:
        n = 300
        pos = rng.uniform(0.0, L, size=(n, 3))
        box = po.Box(upper=(L, L, L), width=w, periodic=per)
        o = po.Oracle(box)
        p = _make(box, pos)
        g = o.exchange_halo(p, n)
        got = sorted((int(p.global_id[n + k]),
                      tuple(Fraction(v) for v in np.array([p.x[n + k], p.y[n + k], p.z[n + k]])
                            - pos[p.global_id[n + k]]))
                     for k in range(g))
        want = []
        Lf, wf = Fraction(L), Fraction(w)
        for i in range(n):
            for d in itertools.product((-1, 0, 1), repeat=3):
                if d == (0, 0, 0) or any(d[a] and not per[a] for a in range(3)):
                    continue
                ok = True
                for a in range(3):
                    img = Fraction(pos[i, a]) - d[a] * Lf
                    if d[a] < 0 and not img < Lf + wf:
                        ok = False
                    if d[a] > 0 and not img > -wf:
                        ok = False
                if ok:
                    want.append((i, tuple(Fraction(float(-d[a] * L)) for a in range(3))))
        # the shift add is exact for these magnitudes only up to rounding, so
        # compare the shift rounded back to the nearest multiple of L
        got_r = sorted((i, tuple(Fraction(round(float(s) / L)) * Lf for s in sh)) for i, sh in got)
        assert got_r == sorted(want)


def test_ghost_copies_and_null_arrays(oracle_lib):
    rng = np.random.default_rng(3)
    n, L = 500, 6.0
    box = po.Box(upper=(L, L, L), width=0.8)
    p = _make(box, rng.uniform(0, L, (n, 3)))
    for name in ("vx", "vy", "vz", "mass", "charge", "fx", "fy", "fz"):
        getattr(p, name)[:n] = rng.standard_normal(n)
    p.type[:n] = rng.integers(0, 5, n)
    o = po.Oracle(box)
    g = o.exchange_halo(p, n)
    own = o.ghost_owners()
    sl = slice(n, n + g)
    for name in ("vx", "vy", "vz", "mass", "charge", "type", "global_id"):
        assert np.array_equal(getattr(p, name)[sl], getattr(p, name)[own])
    assert not p.fx[sl].any() and not p.fy[sl].any() and not p.fz[sl].any()
    # missing velocity arrays (src/halo.cpp:44-53) leave record defaults
    q = _make(box, np.stack([p.x[:n], p.y[:n], p.z[:n]], 1))
    o2 = po.Oracle(box)
    o2.exchange_halo(q, n, omit=("vx", "vy", "vz"))
    assert np.array_equal(q.global_id[n:n + g], p.global_id[sl])
    assert not q.vx.any()


def test_shift_add_maps_negative_zero(oracle_lib):
    """A zero shift is still added, so -0.0 becomes +0.0 in the ghost
    (src/halo.cpp:41-43 adds the shift on every axis)."""
    box = po.Box(lower=(-2.0, -2.0, -2.0), upper=(2.0, 2.0, 2.0), width=0.5)
    p = _make(box, [(1.9, -0.0, 0.0)])
    o = po.Oracle(box)
    assert o.exchange_halo(p, 1) == 1
    assert p.x[1] == 1.9 - 4.0
    assert p.y[1] == 0.0 and not np.signbit(p.y[1]) and np.signbit(p.y[0])


# ---------------------------------------------------------------- reverse
def test_reverse_sum_order_is_ascending_send_slot(oracle_lib):
    """Per owner the ghost values are added in ascending send-slot order
    (src/halo.cpp:1357-1370 with n ascending).  The values cancel
    catastrophically, so only that order gives 4.75 (exact sum: 6.75)."""
    L = 10.0
    box = po.Box(upper=(L, L, L), width=1.0)
    p = _make(box, [(0.2, 0.3, 0.4)])
    o = po.Oracle(box)
    assert o.exchange_halo(p, 1) == 7
    # ghost index of send slot s is 1 + position of block 25-s among
    # the non-empty receive blocks 13,15,16,21,22,24,25
    blocks = [13, 15, 16, 21, 22, 24, 25]
    value = {0: 1e16, 1: 1.0, 3: -1e16, 4: 1.0, 9: 0.5, 10: 0.25, 12: 3.0}
    for s, v in value.items():
        p.fx[1 + blocks.index(25 - s)] = v
    p.fx[0] = 1.0
    o.reverse_halo(p, 1)
    assert p.fx[0] == 4.75
    # Replace: the last writer is the highest send slot
    f = np.zeros(8)
    f[1:8] = p.fx[1:8]
    o.reverse_exchange(1, [po.Field(f, po.REPLACE)])
    assert f[0] == value[12]


def test_reverse_conservation_exact_integers(oracle_lib):
    """Integer-valued forces make the sum order-free: owner total equals its
    value plus the bincount of ghost values by owner."""
    rng = np.random.default_rng(11)
    n, L = 2000, 9.0
    box = po.Box(upper=(L, L, L), width=1.1)
    p = _make(box, rng.uniform(0, L, (n, 3)))
    o = po.Oracle(box)
    g = o.exchange_halo(p, n)
    own = o.ghost_owners()
    for c in ("fx", "fy", "fz"):
        getattr(p, c)[:n + g] = rng.integers(-1000, 1000, n + g).astype(np.float64)
    before = {c: getattr(p, c)[:n].copy() for c in ("fx", "fy", "fz")}
    ghost = {c: getattr(p, c)[n:n + g].copy() for c in ("fx", "fy", "fz")}
    o.reverse_halo(p, n)
    for c in ("fx", "fy", "fz"):
        want = before[c] + np.bincount(own, weights=ghost[c], minlength=n)
        assert np.array_equal(getattr(p, c)[:n], want)
        assert np.array_equal(getattr(p, c)[n:n + g], ghost[c])   # ghosts untouched
    # regression.cpp:66-81: 1.0 per ghost -> owners hold their image count
    cnt = np.zeros(n + g)
    cnt[n:n + g] = 1.0
    o.reverse_exchange(n, [po.Field(cnt, po.SUM)])
    assert np.array_equal(cnt[:n], np.bincount(own, minlength=n).astype(float))
    assert cnt[:n].sum() == g


def test_reverse_min_max_types_and_nan(oracle_lib):
    box = po.Box(upper=(10.0, 10.0, 10.0), width=1.0)
    p = _make(box, [(0.2, 0.3, 0.4), (5.0, 5.0, 0.5)])
    o = po.Oracle(box)
    g = o.exchange_halo(p, 2)
    assert g == 8
    own = o.ghost_owners()
    rng = np.random.default_rng(5)
    for dt in (np.int32, np.int64, np.uint32, np.uint64, np.float32, np.float64):
        vals = rng.integers(1, 100, 10).astype(dt)
        for op, fn in ((po.MIN, np.minimum), (po.MAX, np.maximum)):
            f = vals.copy()
            o.reverse_exchange(2, [po.Field(f, op)])
            for i in range(2):
                want = vals[i]
                for j in np.nonzero(own == i)[0]:
                    want = fn(want, vals[2 + j])
                assert f[i] == want and np.array_equal(f[2:], vals[2:])
    # std::min(dst, v) keeps a NaN owner and ignores a NaN ghost
    f = np.array([np.nan, 3.0] + [1.0] * 8)
    o.reverse_exchange(2, [po.Field(f, po.MIN)])
    assert np.isnan(f[0]) and f[1] == 1.0
    f = np.array([2.0, 3.0] + [np.nan] * 8)
    o.reverse_exchange(2, [po.Field(f, po.MAX)])
    assert f[0] == 2.0 and f[1] == 3.0
    # std::min(0.0, -0.0) keeps the owner's +0.0
    f = np.array([0.0, 0.0] + [-0.0] * 8)
    o.reverse_exchange(2, [po.Field(f, po.MIN)])
    assert not np.signbit(f[0])
    # int32 Sum wraps
    f = np.array([2**31 - 1, 0] + [1] * 8, np.int32)
    o.reverse_exchange(2, [po.Field(f, po.SUM)])
    assert f[0] == np.int32(-(2**31) + 6)


def test_forward_copies_and_adjointness(oracle_lib):
    """forward_exchange.cu:103-167: ghosts hold exact owner copies and
    <fwd(x), y> == <x, rev(y)> within 1e-10 (1 + |.|)."""
    rng = np.random.default_rng(21)
    n, L = 1500, 8.0
    box = po.Box(upper=(L, L, L), width=1.0)
    p = _make(box, rng.uniform(0, L, (n, 3)))
    o = po.Oracle(box)
    g = o.exchange_halo(p, n)
    own = o.ghost_owners()
    feat = np.full((n + g, 7), -99.0)
    feat[:n] = rng.standard_normal((n, 7))
    ff = np.full((n + g, 3), -99.0, np.float32)
    ff[:n] = rng.standard_normal((n, 3)).astype(np.float32)
    o.forward_exchange(n, [po.Field(feat), po.Field(ff)])
    assert np.array_equal(feat[n:], feat[own]) and np.array_equal(ff[n:], ff[own])
    x = rng.standard_normal((n, 7))
    y = rng.standard_normal((g, 7))
    a = np.zeros((n + g, 7))
    a[:n] = x
    o.forward_exchange(n, [po.Field(a)])
    s1 = float((a[n:] * y).sum())
    b = np.zeros((n + g, 7))
    b[n:] = y
    o.reverse_exchange(n, [po.Field(b, po.SUM)])
    s2 = float((x * b[:n]).sum())
    assert abs(s1 - s2) <= 1e-10 * (1 + abs(s1))


# ---------------------------------------------------------------- migrate
def test_migrate_wrap_known_answers(oracle_lib):
    box = po.Box(upper=(32.0, 32.0, 32.0), periodic=(True, True, False))
    x = np.array([-1e-17, 32.5, 5.0, -0.5, 64.25, 31.999999999999996])
    y = x.copy()
    z = x.copy()
    po.lib().dsdxo_migrate_wrap(C_box(box), len(x), po._dp(x), po._dp(y), po._dp(z))
    # x = -1e-17 wraps to exactly upper (no fold-back, src/migration.cpp:106-113)
    assert x[0] == 32.0 and x[1] == 0.5 and x[2] == 5.0 and x[3] == 31.5 and x[4] == 0.25
    assert x[5] == 31.999999999999996
    assert np.array_equal(x, y)
    assert np.array_equal(z, np.array([-1e-17, 32.5, 5.0, -0.5, 64.25, 31.999999999999996]))


def test_migrate_wrap_congruence_exact(oracle_lib):
    """Result lies in [lower, upper] and differs from the input by an integer
    number of extents up to the rounding of the final additions."""
    rng = np.random.default_rng(9)
    lower, upper = -3.7, 12.1
    box = po.Box(lower=(lower,) * 3, upper=(upper,) * 3)
    x = rng.uniform(lower - 40, upper + 40, 4000)
    orig = x.copy()
    y = x.copy()
    z = x.copy()
    po.lib().dsdxo_migrate_wrap(C_box(box), len(x), po._dp(x), po._dp(y), po._dp(z))
    ext = upper - lower
    assert np.all(x >= lower) and np.all(x <= upper)
    k = np.round((orig - x) / ext)
    assert np.all(np.abs(orig - k * ext - x) < 1e-12)
    inside = (orig >= 0) & (orig < ext)
    box0 = po.Box(upper=(ext,) * 3)
    w = orig.copy()
    po.lib().dsdxo_migrate_wrap(C_box(box0), len(w), po._dp(w), po._dp(w.copy()), po._dp(w.copy()))
    assert np.array_equal(w[inside], orig[inside])        # lower == 0: in-box values unchanged


def test_migrate_wrap_rounding_sequence(oracle_lib):
    """The wrap of src/migration.cpp:106-113 at a box whose extent is not a
    power of two and whose lower bound is not zero (lower -3.7, upper 12.1):
    v = x - lower; v -= floor(v / extent) * extent; v + lower, every
    operation rounded on its own.  Expected values come from exact rationals
    with one rounding per step.  The inputs are chosen so that a fused
    multiply-subtract, or subtracting (lower + q * extent) in one go, gives a
    different double for at least one of them, so this test fails for either
    mistake."""
    lower, upper = -3.7, 12.1
    ext = _rn(Fraction(upper) - Fraction(lower))

    def seq(x):
        v = _rn(Fraction(x) - Fraction(lower))
        q = math.floor(_rn(Fraction(v) / Fraction(ext)))
        t = _rn(q * Fraction(ext))
        return _rn(Fraction(_rn(Fraction(v) - Fraction(t))) + Fraction(lower)), v, q

    def fused(x):
        _, v, q = seq(x)
        return _rn(Fraction(_rn(Fraction(v) - q * Fraction(ext))) + Fraction(lower))

    def reordered(x):
        _, v, q = seq(x)
        return _rn(Fraction(x) - Fraction(_rn(Fraction(lower) + Fraction(_rn(q * Fraction(ext))))))

    rng = np.random.default_rng(31)
    xs = list(rng.uniform(lower - 200.0, upper + 200.0, 400)) + [-1e-17, lower, upper, 1000.3, -987.65]
    expected = [seq(x)[0] for x in xs]
    assert any(fused(x) != e for x, e in zip(xs, expected))
    assert any(reordered(x) != e for x, e in zip(xs, expected))
    x = np.array(xs)
    box = po.Box(lower=(lower,) * 3, upper=(upper,) * 3)
    y, z = x.copy(), x.copy()
    po.lib().dsdxo_migrate_wrap(C_box(box), len(x), po._dp(x), po._dp(y), po._dp(z))
    assert x.tolist() == expected
    assert np.array_equal(x, y) and np.array_equal(x, z)


def C_box(box):
    import ctypes
    return ctypes.byref(box.c())
