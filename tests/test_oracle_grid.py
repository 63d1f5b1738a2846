"""Pins the P-rank oracle (oracle/dsdx_oracle_grid.c): the reference examples'
printed outputs on 2/4/8 ranks, closed-form topology, brute-force ownership,
migration invariants and order, a brute-force ghost set and order for every
rank, conservation and adjointness of the reverse/forward exchanges, and the
1x1x1 loopback oracle when the unsplit-axis skip is lifted."""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as po


def _one_per_rank(g, offset):
    parts = [po.Particles(64) for _ in range(g.ranks)]
    for r, p in enumerate(parts):
        lo, _ = g.local(r)
        p.x[0], p.y[0], p.z[0] = lo[0] + offset, lo[1] + offset, lo[2] + offset
        p.global_id[0] = r
        p.mass[0] = 1.0
    return parts, [1] * g.ranks


@pytest.mark.parametrize("dims,ghosts", [((1, 1, 1), 0), ((2, 1, 1), 1), ((2, 2, 1), 3),
                                         ((2, 2, 2), 7)])
def test_getting_started_example(dims, ghosts):
    """examples/getting_started: L=8, w=0.25, one particle per rank at its
    subdomain's lower corner + 0.05; 'rank r ghosts=G returned_fx=G'
    (examples/README.md:10-16 for 2 ranks: ghosts=1 returned_fx=1)."""
    g = po.Grid(po.Box(upper=(8.0, 8.0, 8.0), width=0.25), dims)
    parts, owned = _one_per_rank(g, 0.05)
    g.migrate(parts, owned)
    gh = g.exchange_halo(parts, owned)
    assert gh == [ghosts] * g.ranks
    for r, p in enumerate(parts):
        p.fx[:] = 0.0
        p.fx[owned[r]:owned[r] + gh[r]] = 1.0
    g.reverse_halo(parts, owned)
    assert [p.fx[0] for p in parts] == [float(ghosts)] * g.ranks


def test_domain_decomposition_and_halo_examples():
    """examples/domain_decomposition ('rank 0 grid 2x1x1 coordinate 0,0,0
    neighbors 2') and examples/halo_exchange ('rank 0 received 1 ghosts')."""
    g = po.Grid(po.Box(upper=(10.0, 10.0, 10.0), width=0.5), (2, 1, 1))
    assert g.coord(0) == (0, 0, 0) and len(g.neighbor_list(0)) == 2
    parts, owned = _one_per_rank(g, 0.1)
    assert g.exchange_halo(parts, owned) == [1, 1]


def test_particle_migration_example():
    """examples/particle_migration on 2 ranks: x = 10.25 on rank 0 and -0.25
    on rank 1 wrap to 0.25 and 9.75, each owned by its own rank ('rank r now
    owns 1 particles')."""
    g = po.Grid(po.Box(upper=(10.0, 10.0, 10.0)), (2, 1, 1))
    parts = [po.Particles(8) for _ in range(2)]
    parts[0].x[0], parts[1].x[0] = 10.25, -0.25
    for r, p in enumerate(parts):
        p.y[0] = p.z[0] = 5.0
        p.global_id[0] = r
    owned = [1, 1]
    sent, recv = g.migrate(parts, owned)
    assert owned == [1, 1] and sent == [0, 0] and recv == [0, 0]
    assert parts[0].x[0] == 0.25 and parts[1].x[0] == 9.75


def test_record_on_the_upper_face_belongs_to_the_first_subdomain():
    """x = -1e-17 wraps to exactly 10.0 (no fold-back, wrapped_record), and
    migrate asks owner_rank for that position, which wraps it again to 0.0
    (src/migration.cpp:83-87, :440): the particle is stored at 10.0 but owned
    by the first subdomain.  A single wrap would give it to the last one."""
    g = po.Grid(po.Box(upper=(10.0, 10.0, 10.0)), (2, 1, 1))
    parts = [po.Particles(8) for _ in range(2)]
    parts[1].x[0], parts[1].y[0], parts[1].z[0] = -1e-17, 5.0, 5.0
    parts[1].global_id[0] = 7
    owned = [0, 1]
    g.migrate(parts, owned)
    assert owned == [1, 0]
    assert parts[0].x[0] == 10.0 and parts[0].global_id[0] == 7
    assert g.owner_rank(10.0, 5.0, 5.0) == 0


@pytest.mark.parametrize("dims", [(2, 1, 1), (1, 3, 1), (2, 2, 2), (3, 2, 1), (4, 1, 2)])
@pytest.mark.parametrize("periodic", [(True, True, True), (True, False, True)])
def test_topology_closed_form(dims, periodic):
    box = po.Box(lower=(-1.0, 0.5, 2.0), upper=(7.0, 9.5, 5.0), periodic=periodic, width=0.3)
    g = po.Grid(box, dims)
    assert g.ranks == int(np.prod(dims))
    for a in range(3):
        ext = box.upper[a] - box.lower[a]
        assert g.cuts(a) == [box.lower[a] + ext * i / dims[a] for i in range(dims[a] + 1)]
    for c in itertools.product(*[range(d) for d in dims]):
        r = (c[0] * dims[1] + c[1]) * dims[2] + c[2]
        assert g.coord(r) == c
        lo, hi = g.local(r)
        for a in range(3):
            assert lo[a] == g.cuts(a)[c[a]] and hi[a] == g.cuts(a)[c[a] + 1]
        expect = []
        for d in itertools.product((-1, 0, 1), repeat=3):
            if d == (0, 0, 0) or any(d[a] and dims[a] == 1 for a in range(3)):
                continue
            nc, shift, out = list(c), [0.0] * 3, False
            for a in range(3):
                nc[a] += d[a]
                if not 0 <= nc[a] < dims[a]:
                    if not periodic[a]:
                        out = True
                        break
                    ext = box.upper[a] - box.lower[a]
                    shift[a] = ext if nc[a] < 0 else -ext
                    nc[a] %= dims[a]
            rank = -1 if out else (nc[0] * dims[1] + nc[1]) * dims[2] + nc[2]
            slot = (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1)
            slot = slot if slot < 13 else slot - 1
            expect.append((d, slot, rank, tuple(shift) if not out else (0.0, 0.0, 0.0)))
        assert g.neighbor_list(r) == expect


def test_owner_rank_brute_force():
    box = po.Box(upper=(6.0, 6.0, 6.0), periodic=(True, True, False))
    g = po.Grid(box, (3, 2, 2))
    rng = np.random.default_rng(3)
    pts = rng.uniform(-7.0, 13.0, (4000, 3))
    # points exactly on the cut planes of every axis (upper_bound ties)
    cuts = np.array(sorted(set(g.cuts(0) + g.cuts(1) + g.cuts(2))))
    pts[:len(cuts) ** 3] = np.array(list(itertools.product(cuts, repeat=3)))
    for x, y, z in pts:
        v = [x, y, z]
        for a in range(2):
            t = v[a] - 0.0
            v[a] = t - np.floor(t / 6.0) * 6.0
        c = []
        for a in range(3):
            cuts = g.cuts(a)
            k = sum(1 for cut in cuts if not v[a] < cut)      # upper_bound index
            c.append(min(max(k - 1, 0), g.dims[a] - 1))
        assert g.owner_rank(x, y, z) == (c[0] * 2 + c[1]) * 2 + c[2]


def _random_grid_particles(g, n_per_rank, spread, seed, fields=False):
    rng = np.random.default_rng(seed)
    cap = 8 * n_per_rank + 256
    parts, owned, field_sets = [], [], []
    gid = 0
    L = np.array(g.box.upper) - np.array(g.box.lower)
    for r in range(g.ranks):
        p = po.Particles(cap)
        lo, hi = np.array(g.local(r)[0]), np.array(g.local(r)[1])
        pos = rng.uniform(lo - spread * L, hi + spread * L, (n_per_rank, 3))
        p.x[:n_per_rank], p.y[:n_per_rank], p.z[:n_per_rank] = pos.T
        for k in ("vx", "vy", "vz", "fx", "fy", "fz", "charge"):
            getattr(p, k)[:n_per_rank] = rng.standard_normal(n_per_rank)
        p.mass[:n_per_rank] = rng.uniform(0.5, 2.0, n_per_rank)
        p.type[:n_per_rank] = rng.integers(0, 4, n_per_rank)
        p.global_id[:n_per_rank] = np.arange(gid, gid + n_per_rank)
        gid += n_per_rank
        parts.append(p)
        owned.append(n_per_rank)
        if fields:
            field_sets.append([po.Field(rng.standard_normal(cap)),
                               po.Field(rng.standard_normal((cap, 3)))])
    return parts, owned, field_sets


@pytest.mark.parametrize("dims,periodic", [((2, 2, 2), (True, True, True)),
                                           ((3, 1, 2), (True, True, False)),
                                           ((4, 1, 1), (True, True, True))])
def test_migrate_invariants_and_order(dims, periodic):
    box = po.Box(upper=(9.0, 6.0, 7.5), periodic=periodic, width=0.4)
    g = po.Grid(box, dims)
    parts, owned, fields = _random_grid_particles(g, 300, 0.3, 7, fields=True)
    before = [{k: getattr(p, k)[:owned[r]].copy() for k in po.Particles.F64 + po.Particles.I32}
              for r, p in enumerate(parts)]
    fbefore = [[f.data[:owned[r]].copy() for f in fields[r]] for r in range(g.ranks)]
    old_owned = list(owned)
    g.migrate(parts, owned, fields)
    # expected placement from first principles: wrap, owner, retained first
    recs = []
    for r in range(g.ranks):
        for i in range(old_owned[r]):
            v = [before[r]["x"][i], before[r]["y"][i], before[r]["z"][i]]
            for a in range(3):
                if periodic[a]:
                    ext = box.upper[a] - box.lower[a]
                    t = v[a] - box.lower[a]
                    t = t - np.floor(t / ext) * ext
                    v[a] = t + box.lower[a]
            recs.append((r, i, g.owner_rank(*v), v))
    for q in range(g.ranks):
        order = [(r, i, v) for r, i, d, v in recs if d == q and r == q] + \
                [(r, i, v) for s in range(g.ranks) if s != q for r, i, d, v in recs
                 if d == q and r == s]
        assert owned[q] == len(order)
        p = parts[q]
        for w, (r, i, v) in enumerate(order):
            assert p.global_id[w] == before[r]["global_id"][i]
            assert (p.x[w], p.y[w], p.z[w]) == tuple(v)
            assert p.vx[w] == before[r]["vx"][i] and p.fz[w] == before[r]["fz"][i]
            np.testing.assert_array_equal(fields[q][1].data[w], fbefore[r][1][i])
            lo, hi = g.local(q)
            for a in range(3):
                if periodic[a]:
                    assert lo[a] <= v[a] <= hi[a]
    ids = np.sort(np.concatenate([parts[q].global_id[:owned[q]] for q in range(g.ranks)]))
    np.testing.assert_array_equal(ids, np.arange(sum(old_owned)))


@pytest.mark.parametrize("dims,periodic,corners", [((2, 2, 2), (True, True, True), True),
                                                   ((3, 2, 1), (True, False, True), True),
                                                   ((2, 3, 2), (True, True, True), False)])
def test_halo_ghost_set_and_order_brute_force(dims, periodic, corners):
    box = po.Box(upper=(8.0, 9.0, 7.0), periodic=periodic, width=0.9, edges_and_corners=corners)
    g = po.Grid(box, dims)
    parts, owned, _ = _random_grid_particles(g, 400, 0.0, 11)
    g.migrate(parts, owned)
    snap = [{k: getattr(p, k)[:owned[r]].copy() for k in ("x", "y", "z", "global_id", "vx")}
            for r, p in enumerate(parts)]
    gh = g.exchange_halo(parts, owned)

    def sends(s, d):
        lo, hi = g.local(s)
        sel = []
        for i in range(owned[s]):
            v = (snap[s]["x"][i], snap[s]["y"][i], snap[s]["z"][i])
            ok = True
            for a in range(3):
                if d[a] < 0 and not v[a] < lo[a] + box.width:
                    ok = False
                if d[a] > 0 and not v[a] > hi[a] - box.width:
                    ok = False
            if ok:
                sel.append(i)
        return sel

    for q in range(g.ranks):
        expect_id, expect_x = [], []
        for d, slot, s, shift in g.neighbor_list(q):
            if s < 0:
                continue
            back = tuple(-x for x in d)
            entry = [e for e in g.neighbor_list(s) if e[0] == back][0]
            for i in sends(s, back):
                expect_id.append(snap[s]["global_id"][i])
                expect_x.append(snap[s]["x"][i] + entry[3][0])
        n = owned[q]
        assert gh[q] == len(expect_id)
        np.testing.assert_array_equal(parts[q].global_id[n:n + gh[q]], expect_id)
        np.testing.assert_array_equal(parts[q].x[n:n + gh[q]], expect_x)
        assert (parts[q].fx[n:n + gh[q]] == 0.0).all()


def test_loopback_equals_single_box_oracle():
    box = po.Box(upper=(12.0, 12.0, 12.0), width=2.0)
    g = po.Grid(box, (1, 1, 1), lift_unsplit=True)
    rng = np.random.default_rng(5)
    n = 3000
    a, b = po.Particles(4 * n), po.Particles(4 * n)
    vals = {k: rng.uniform(-1, 13, n) if k in "xyz" else rng.standard_normal(n)
            for k in po.Particles.F64}
    for k, v in vals.items():
        getattr(a, k)[:n] = v
        getattr(b, k)[:n] = v
    a.global_id[:n] = b.global_id[:n] = np.arange(n)
    o = po.Oracle(box)
    o.migrate(a, n)
    ga = o.exchange_halo(a, n)
    owned = [n]
    g.migrate([b], owned)
    gb = g.exchange_halo([b], owned)
    assert [ga] == gb and owned == [n]
    for k in po.Particles.F64 + po.Particles.I32:
        np.testing.assert_array_equal(getattr(a, k)[:n + ga], getattr(b, k)[:n + ga])
    f = rng.standard_normal((3, n + ga))
    for p in (a, b):
        p.fx[:n + ga], p.fy[:n + ga], p.fz[:n + ga] = f
    o.reverse_halo(a, n)
    g.reverse_halo([b], [n])
    for k in ("fx", "fy", "fz"):
        np.testing.assert_array_equal(getattr(a, k)[:n], getattr(b, k)[:n])


def test_reverse_conservation_and_forward_adjoint():
    box = po.Box(upper=(8.0, 8.0, 8.0), periodic=(True, True, False), width=1.1)
    g = po.Grid(box, (2, 2, 2))
    parts, owned, _ = _random_grid_particles(g, 500, 0.0, 17)
    g.migrate(parts, owned)
    gh = g.exchange_halo(parts, owned)
    rng = np.random.default_rng(18)
    total_before = 0
    for r, p in enumerate(parts):
        p.fx[:owned[r] + gh[r]] = rng.integers(-1000, 1000, owned[r] + gh[r]).astype(np.float64)
        total_before += p.fx[:owned[r] + gh[r]].sum()
    g.reverse_halo(parts, owned)
    total_after = sum(p.fx[:owned[r]].sum() for r, p in enumerate(parts))
    assert total_after == total_before
    # forward / reverse(Sum) adjointness: <F x, y> == <x, R y>
    xs = [rng.standard_normal(p.capacity) for p in parts]
    ys = [rng.standard_normal(p.capacity) for p in parts]
    fx = [[po.Field(x.copy())] for x in xs]
    for r in range(g.ranks):
        fx[r][0].data[owned[r]:] = 0.0
    g.forward_exchange(owned, fx)
    lhs = sum(float(fx[r][0].data[owned[r]:owned[r] + gh[r]] @ ys[r][owned[r]:owned[r] + gh[r]])
              for r in range(g.ranks))
    ry = [[po.Field(np.where(np.arange(p.capacity) >= owned[r], ys[r], 0.0))]
          for r, p in enumerate(parts)]
    g.reverse_exchange(owned, ry)
    rhs = sum(float(xs[r][:owned[r]] @ ry[r][0].data[:owned[r]]) for r in range(g.ranks))
    assert abs(lhs - rhs) <= 1e-10 * (1 + abs(lhs))


def test_reference_run_on_eight_ranks_bit_exact():
    """The fixture tests/golden/ref_grid_2x2x2.npz holds what the reference
    itself (its unmodified Host library on 8 MPI ranks, a 2x2x2 periodic
    grid; tests/golden/make_ref_fixtures.py) held on every rank after each
    operation.  The box does not start at 0, and its extents are not powers of
    two.  3000 particles start on the wrong ranks, and 10 % lie past each face.
    The oracle, on the same inputs, must reproduce it bit for bit over two
    steps:
    - after migrate: owned ids in order, positions (wrapped), velocities and
      the registered field;
    - after exchange_halo: ghost ids in order, positions, velocities, masses,
      types and the field;
    - after reverse_halo: the owned forces;
    - after reverse_exchange(Sum) of the field: the owned values."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_grid_2x2x2.npz")
    with np.load(path) as z:
        ref = {k: z[k] for k in z.files}
    box = po.Box(lower=tuple(ref["box_lower"]), upper=tuple(ref["box_upper"]), width=float(ref["box_width"][0]))
    g = po.Grid(box, tuple(int(v) for v in ref["dims"]))
    cap = 4 * 3000
    parts = [po.Particles(cap) for _ in range(g.ranks)]
    aux = [np.zeros((cap, 3)) for _ in range(g.ranks)]
    owned = []
    for r, p in enumerate(parts):
        n = len(ref[f"r{r}_s0_start_global_id"])
        for k in ("global_id", "type", "x", "y", "z", "vx", "vy", "vz", "mass", "charge"):
            getattr(p, k)[:n] = ref[f"r{r}_s0_start_{k}"]
        aux[r][:n] = ref[f"r{r}_s0_start_aux"].reshape(n, 3)
        owned.append(n)

    def same(a, b, what):
        a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
        assert a.shape == b.shape, what
        if a.dtype == np.float64:
            a, b = a.view(np.uint64), b.view(np.uint64)
        assert np.array_equal(a, b), what

    for s in range(2):
        fields = [[po.Field(aux[r])] for r in range(g.ranks)]
        g.migrate(parts, owned, fields)
        for r, p in enumerate(parts):
            n = owned[r]
            pre = f"r{r}_s{s}_"
            same(p.global_id[:n], ref[pre + "migrate_global_id"], f"step {s} rank {r} owned ids")
            for k in ("x", "y", "z", "vx"):
                same(getattr(p, k)[:n], ref[pre + "migrate_" + k], f"step {s} rank {r} migrate {k}")
            same(aux[r][:n].ravel(), ref[pre + "migrate_aux"], f"step {s} rank {r} migrate aux")
            for k in ("x", "y", "z"):
                getattr(p, k)[:n] = ref[pre + "ties_" + k]
        ghosts = g.exchange_halo(parts, owned, fields)
        for r, p in enumerate(parts):
            n, gh = owned[r], ghosts[r]
            pre = f"r{r}_s{s}_"
            same(p.global_id[n:n + gh], ref[pre + "halo_global_id"], f"step {s} rank {r} ghost ids")
            for k in ("x", "y", "z", "vz", "mass", "type"):
                same(getattr(p, k)[n:n + gh], ref[pre + "halo_" + k], f"step {s} rank {r} ghost {k}")
            same(aux[r][n:n + gh].ravel(), ref[pre + "halo_aux"], f"step {s} rank {r} ghost aux")
            for k in ("fx", "fy", "fz"):
                getattr(p, k)[:n + gh] = ref[pre + "forces_" + k]
        g.reverse_halo(parts, owned)
        for r, p in enumerate(parts):
            n, gh = owned[r], ghosts[r]
            pre = f"r{r}_s{s}_"
            for k in ("fx", "fy", "fz"):
                same(getattr(p, k)[:n], ref[pre + "reverse_" + k], f"step {s} rank {r} reverse {k}")
            aux[r][n:n + gh] = ref[pre + "aux_ghosts"].reshape(gh, 3)
        g.reverse_exchange(owned, [[po.Field(aux[r])] for r in range(g.ranks)])
        for r, p in enumerate(parts):
            n = owned[r]
            pre = f"r{r}_s{s}_"
            same(aux[r][:n].ravel(), ref[pre + "aux_reduced"], f"step {s} rank {r} reduced aux")
            for k in ("x", "y", "z"):
                getattr(p, k)[:n] = ref[pre + "displaced_" + k]
