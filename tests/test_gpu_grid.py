"""Process grids on one GPU (dsdx_grid_*): every subdomain's migration, halo
and reverse sums bit-exact with the oracle's Host path on the same number of
ranks (oracle/dsdx_oracle_grid.c), over several steps; and the reference
example getting_started on 2, 4 and 8 subdomains."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import F64, I32, assert_bits_equal, host, say

from tests.gpu_util import ghost_writes  # noqa: F401  (autouse: both ghost-write kernels)
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


def _load(torch, grid, parts, owned):
    """Oracle per-rank arrays -> the grid's current device storage."""
    for r, p in enumerate(parts):
        sl = grid.slots(r, owned[r])
        for k in F64 + I32:
            getattr(grid.particles, k)[sl] = torch.from_numpy(getattr(p, k)[:owned[r]]).cuda()
    grid.set_owned(owned)


def _compare(grid, parts, owned, ghosts, what, keys=F64 + I32):
    dev = {k: host(getattr(grid.particles, k)) for k in keys}
    for r, p in enumerate(parts):
        sl = grid.slots(r, owned[r] + ghosts[r])
        for k in keys:
            assert_bits_equal(dev[k][sl], getattr(p, k)[:owned[r] + ghosts[r]], f"{what} rank {r} {k}")


@pytest.mark.parametrize("dims,ghosts", [((2, 1, 1), 1), ((2, 2, 1), 3), ((2, 2, 2), 7)])
def test_getting_started_on_subdomains(gpu, dims, ghosts):
    """examples/getting_started: one particle per rank at its subdomain's
    lower corner + 0.05 (L = 8, w = 0.25): 'ghosts=G returned_fx=G'."""
    torch = gpu
    from dsdecompx_amd.grid import SubdomainGrid
    from dsdecompx_amd.halo import BoxConfig
    box = BoxConfig(upper=(8.0, 8.0, 8.0), width=0.25)
    grid = SubdomainGrid(box, dims, 64)
    og = po.Grid(po.Box(upper=(8.0, 8.0, 8.0), width=0.25), dims)
    p = grid.particles
    for r in range(grid.ranks):
        lo, _ = og.local(r)
        j = r * grid.capacity
        p.x[j], p.y[j], p.z[j] = lo[0] + 0.05, lo[1] + 0.05, lo[2] + 0.05
        p.global_id[j] = r
        p.mass[j] = 1.0
    grid.set_owned([1] * grid.ranks)
    assert grid.migrate() == [1] * grid.ranks
    assert grid.exchange_halo() == [ghosts] * grid.ranks
    p = grid.particles
    p.fx.zero_()
    for r in range(grid.ranks):
        p.fx[grid.slots(r, 1 + ghosts)][1:] = 1.0
    grid.reverse_halo()
    torch.cuda.synchronize()
    assert [float(p.fx[r * grid.capacity]) for r in range(grid.ranks)] == [float(ghosts)] * grid.ranks


CASES = [
    # dims, upper, periodic, width, edges_and_corners, particles per rank
    ((2, 2, 2), (16.0, 16.0, 16.0), (True, True, True), 1.4, True, 3000),
    ((3, 2, 1), (21.0, 14.0, 9.0), (True, False, True), 1.1, True, 2500),
    ((4, 1, 1), (24.0, 7.0, 7.0), (True, True, True), 0.9, True, 2000),
    ((2, 3, 2), (12.0, 15.0, 10.0), (True, True, False), 1.6, False, 1500),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_steps_bit_exact_with_oracle(gpu, case):
    torch = gpu
    from dsdecompx_amd.grid import SubdomainGrid
    from dsdecompx_amd.halo import BoxConfig
    dims, upper, periodic, w, corners, n = case
    say("grid", dims)
    obox = po.Box(upper=upper, periodic=periodic, width=w, edges_and_corners=corners)
    og = po.Grid(obox, dims)
    cap = 4 * n + 512
    grid = SubdomainGrid(BoxConfig(upper=upper, periodic=periodic, width=w,
                                   edges_and_corners=corners), dims, cap)
    rng = np.random.default_rng(41)
    parts, owned = [], []
    gid = 0
    L = np.array(upper)
    for r in range(og.ranks):
        p = po.Particles(cap)
        lo, hi = np.array(og.local(r)[0]), np.array(og.local(r)[1])
        # some particles start outside their subdomain (and outside the box)
        pos = rng.uniform(lo - 0.05 * L, hi + 0.05 * L, (n, 3))
        for a in range(3):
            if not periodic[a]:
                pos[:, a] = np.clip(pos[:, a], 0.0, upper[a] * (1 - 1e-12))
        p.x[:n], p.y[:n], p.z[:n] = pos.T
        for k in ("vx", "vy", "vz", "fx", "fy", "fz", "charge"):
            getattr(p, k)[:n] = rng.standard_normal(n)
        p.mass[:n] = rng.uniform(0.5, 2.0, n)
        p.type[:n] = rng.integers(0, 4, n)
        p.global_id[:n] = np.arange(gid, gid + n)
        gid += n
        parts.append(p)
        owned.append(n)
    _load(torch, grid, parts, owned)
    for step in range(3):
        if step:
            # displace every owned particle on both sides
            disp = [rng.normal(0, 0.3, (owned[r], 3)) for r in range(og.ranks)]
            for r, p in enumerate(parts):
                for a, k in enumerate(("x", "y", "z")):
                    getattr(p, k)[:owned[r]] += disp[r][:, a]
                    if not periodic[a]:
                        v = getattr(p, k)[:owned[r]]
                        np.clip(v, 0.0, upper[a] * (1 - 1e-12), out=v)
                    dv = torch.from_numpy(getattr(p, k)[:owned[r]].copy()).cuda()
                    getattr(grid.particles, k)[grid.slots(r, owned[r])] = dv
        og.migrate(parts, owned)
        assert grid.migrate() == owned, f"step {step} owned counts"
        _compare(grid, parts, owned, [0] * og.ranks, f"step {step} migrate")
        ghosts = og.exchange_halo(parts, owned)
        assert grid.exchange_halo() == ghosts, f"step {step} ghost counts"
        _compare(grid, parts, owned, ghosts, f"step {step} halo")
        # distinct boundary owners: owned particles with an image anywhere
        imaged = set()
        for r, p in enumerate(parts):
            imaged.update(p.global_id[owned[r]:owned[r] + ghosts[r]].tolist())
        owners = [len(imaged.intersection(p.global_id[:owned[r]].tolist())) for r, p in enumerate(parts)]
        assert grid.counts()[2] == owners, f"step {step} owners"
        f = rng.uniform(-1, 1, (og.ranks, 3, cap))
        for r, p in enumerate(parts):
            m = owned[r] + ghosts[r]
            for a, k in enumerate(("fx", "fy", "fz")):
                getattr(p, k)[:m] = f[r, a, :m]
                getattr(grid.particles, k)[grid.slots(r, m)] = torch.from_numpy(f[r, a, :m]).cuda()
        og.reverse_halo(parts, owned)
        grid.reverse_halo()
        _compare(grid, parts, owned, [0] * og.ranks, f"step {step} reverse", ("fx", "fy", "fz"))


def _random_setup(torch, dims, upper, periodic, w, corners, n, seed, cap=None):
    from dsdecompx_amd.grid import SubdomainGrid
    from dsdecompx_amd.halo import BoxConfig
    og = po.Grid(po.Box(upper=upper, periodic=periodic, width=w, edges_and_corners=corners), dims)
    cap = cap or 4 * n + 512
    grid = SubdomainGrid(BoxConfig(upper=upper, periodic=periodic, width=w,
                                   edges_and_corners=corners), dims, cap)
    rng = np.random.default_rng(seed)
    parts, owned = [], []
    L = np.array(upper)
    for r in range(og.ranks):
        p = po.Particles(cap)
        lo, hi = np.array(og.local(r)[0]), np.array(og.local(r)[1])
        pos = rng.uniform(lo - 0.05 * L, hi + 0.05 * L, (n, 3))
        for a in range(3):
            if not periodic[a]:
                pos[:, a] = np.clip(pos[:, a], 0.0, upper[a] * (1 - 1e-12))
        p.x[:n], p.y[:n], p.z[:n] = pos.T
        for k in ("vx", "vy", "vz", "fx", "fy", "fz", "charge"):
            getattr(p, k)[:n] = rng.standard_normal(n)
        p.mass[:n] = rng.uniform(0.5, 2.0, n)
        p.type[:n] = rng.integers(0, 4, n)
        p.global_id[:n] = np.arange(r * n, (r + 1) * n)
        parts.append(p)
        owned.append(n)
    _load(torch, grid, parts, owned)
    return og, grid, parts, owned, rng


# name, dtype, components, op: one field of each type and reduction
FIELDS = [("q64", np.float64, 1, po.SUM), ("v32", np.float32, 3, po.MIN),
          ("i32", np.int32, 2, po.MAX), ("u64", np.uint64, 1, po.REPLACE),
          ("i64", np.int64, 1, po.SUM), ("u32", np.uint32, 1, po.SUM)]


def _field_values(rng, dtype, shape):
    if np.issubdtype(dtype, np.floating):
        return rng.standard_normal(shape).astype(dtype)
    info = np.iinfo(dtype)
    return rng.integers(max(info.min, -2**40), min(info.max, 2**40), shape, dtype=dtype, endpoint=True)


def test_fields_migrate_halo_reverse_forward(gpu):
    """Registered fields of every type on a 2x2x2 grid: carried by migration
    and into the ghosts by the halo build, reduced by reverse_exchange with
    each ReduceOp, copied by forward_exchange -- bit-exact with the oracle."""
    torch = gpu
    say("grid fields")
    og, grid, parts, owned, rng = _random_setup(torch, (2, 2, 2), (12.0, 12.0, 12.0),
                                                (True, True, True), 1.2, True, 2000, 5)
    cap = grid.capacity
    of = [[] for _ in range(og.ranks)]
    for name, dt, comps, op in FIELDS:
        tdt = {np.float64: torch.float64, np.float32: torch.float32, np.int32: torch.int32,
               np.uint64: torch.uint64, np.int64: torch.int64, np.uint32: torch.uint32}[dt]
        grid.add_field(name, tdt, comps, op)
        shape = (cap,) if comps == 1 else (cap, comps)
        for r in range(og.ranks):
            a = np.zeros(shape, dt)
            a[:owned[r]] = _field_values(rng, dt, (owned[r],) + shape[1:])
            of[r].append(po.Field(a, op))
            grid.field(name)[grid.slots(r, owned[r])] = torch.from_numpy(a[:owned[r]]).cuda()

    def compare(what, counts):
        for f, (name, *_rest) in enumerate(FIELDS):
            dev = host(grid.field(name))
            for r in range(og.ranks):
                sl = grid.slots(r, counts[r])
                assert_bits_equal(dev[sl], of[r][f].data[:counts[r]], f"{what} rank {r} {name}")

    og.migrate(parts, owned, of)
    assert grid.migrate() == owned
    compare("migrate", owned)
    ghosts = og.exchange_halo(parts, owned, of)
    assert grid.exchange_halo() == ghosts
    compare("halo", [o + g for o, g in zip(owned, ghosts)])
    _compare(grid, parts, owned, ghosts, "halo records")
    # reverse: fresh values in every ghost slot
    for f, (name, dt, comps, op) in enumerate(FIELDS):
        for r in range(og.ranks):
            m = owned[r] + ghosts[r]
            a = of[r][f].data
            a[owned[r]:m] = _field_values(rng, dt, (ghosts[r],) + a.shape[1:])
            grid.field(name)[grid.slots(r, m)] = torch.from_numpy(a[:m]).cuda()
    og.reverse_exchange(owned, of)
    grid.reverse_exchange()
    compare("reverse", owned)
    # forward: fresh owner values
    for f, (name, dt, comps, op) in enumerate(FIELDS):
        for r in range(og.ranks):
            a = of[r][f].data
            a[:owned[r]] = _field_values(rng, dt, (owned[r],) + a.shape[1:])
            grid.field(name)[grid.slots(r, owned[r])] = torch.from_numpy(a[:owned[r]]).cuda()
    og.forward_exchange(owned, of)
    grid.forward_exchange()
    compare("forward", [o + g for o, g in zip(owned, ghosts)])


@pytest.mark.parametrize("positions_only", [False, True])
def test_refresh_rewrites_ghosts_from_owners(gpu, positions_only):
    """A reused plan on a 3x2x1 grid (y non-periodic): after the owners move
    and change, every ghost equals its owner (found by global id, in whichever
    subdomain) with the build's periodic shift added to the new position;
    positions_only leaves the other members and zeroes the forces."""
    torch = gpu
    say("grid refresh", positions_only)
    upper, periodic = (18.0, 12.0, 9.0), (True, False, True)
    og, grid, parts, owned, rng = _random_setup(torch, (3, 2, 1), upper, periodic, 1.0, True,
                                                1500, 9)
    og.migrate(parts, owned)
    grid.migrate()
    ghosts = grid.exchange_halo()
    p = grid.particles
    before = {k: host(getattr(p, k)).copy() for k in F64 + I32}
    # owners move (well inside the box) and their other members change
    n_all = grid.ranks * grid.capacity
    owned_mask = np.zeros(n_all, bool)
    for r in range(grid.ranks):
        owned_mask[grid.slots(r, owned[r])] = True
    after = {k: v.copy() for k, v in before.items()}
    for k in ("x", "y", "z"):
        after[k][owned_mask] += rng.normal(0, 0.01, owned_mask.sum())
    for k in ("vx", "mass"):
        after[k][owned_mask] += 1.0
    for k in ("x", "y", "z", "vx", "mass"):
        getattr(p, k).copy_(torch.from_numpy(after[k]))
    p.fx.fill_(3.0)
    grid.refresh(positions_only=positions_only)
    dev = {k: host(getattr(p, k)) for k in F64 + I32}
    where = {int(before["global_id"][j]): j for j in np.flatnonzero(owned_mask)}
    L = np.array(upper)
    for r in range(grid.ranks):
        for j in range(r * grid.capacity + owned[r], r * grid.capacity + owned[r] + ghosts[r]):
            o = where[int(before["global_id"][j])]
            for a, k in enumerate(("x", "y", "z")):
                shift = np.round((before[k][j] - before[k][o]) / L[a]) * L[a]
                assert after[k][o] + shift == dev[k][j], (r, j, k)
            for k in ("vx", "vy", "vz", "mass", "charge", "type", "global_id"):
                expect = before[k][j] if positions_only else after[k][o]
                assert expect == dev[k][j], (r, j, k)
            for k in ("fx", "fy", "fz"):
                assert dev[k][j] == 0.0, (r, j, k)


def test_overflow_is_reported_and_writes_nothing(gpu):
    """A migration that would overfill a subdomain, and a halo that would,
    raise CapacityError; the target storage and the ghost slots keep their
    contents, and the grid still builds the oracle's halo afterwards."""
    torch = gpu
    from dsdecompx_amd._abi import CapacityError
    from dsdecompx_amd.grid import SubdomainGrid
    from dsdecompx_amd.halo import BoxConfig
    say("grid overflow")
    upper, cap = (8.0, 8.0, 8.0), 160
    og = po.Grid(po.Box(upper=upper, width=1.5), (2, 1, 1))
    grid = SubdomainGrid(BoxConfig(upper=upper, width=1.5), (2, 1, 1), cap)
    rng = np.random.default_rng(3)
    parts = []
    owned = [120, 60]
    for r, n in enumerate(owned):
        p = po.Particles(cap)
        # every particle of rank 0 belongs to rank 1: 180 > 160 after migration
        p.x[:n] = rng.uniform(4.0, 8.0, n)
        p.y[:n], p.z[:n] = rng.uniform(0, 8, n), rng.uniform(0, 8, n)
        p.global_id[:n] = np.arange(n) + 1000 * r
        parts.append(p)
    _load(torch, grid, parts, owned)
    spare = grid._storage[1]
    spare.x.fill_(-7.0)
    with pytest.raises(CapacityError):
        grid.migrate()
    assert grid._current == 0 and grid.counts()[0] == owned
    assert bool((host(spare.x) == -7.0).all())
    # halo: 120 + 60 owned near the x faces of 2 subdomains cannot hold the ghosts
    grid.particles.x[grid.slots(0, 120)] = torch.from_numpy(rng.uniform(0.0, 0.5, 120)).cuda()
    grid.particles.x[grid.slots(1, 60)] = torch.from_numpy(rng.uniform(7.5, 8.0, 60)).cuda()
    grid.particles.y.fill_(4.0)
    grid.particles.z.fill_(4.0)
    grid.particles.x[grid.slots(0, cap)][120:] = -9.0
    grid.particles.x[grid.slots(1, cap)][60:] = -9.0
    with pytest.raises(CapacityError):
        grid.exchange_halo()
    xs = host(grid.particles.x)
    assert (xs[120:cap] == -9.0).all() and (xs[cap + 60:] == -9.0).all()
    # back to a state that fits (rank 1 interior, about half of rank 0 near
    # its +x face): the oracle's halo, bit-exact
    parts[0].x[:120] = rng.uniform(2.0, 3.0, 120)
    parts[1].x[:60] = rng.uniform(5.6, 6.4, 60)
    for r, p in enumerate(parts):
        p.y[:owned[r]] = 4.0
        p.z[:owned[r]] = 4.0
        grid.particles.x[grid.slots(r, owned[r])] = torch.from_numpy(p.x[:owned[r]].copy()).cuda()
    ghosts = og.exchange_halo(parts, owned)
    assert grid.exchange_halo() == ghosts
    _compare(grid, parts, owned, ghosts, "after overflow")


def test_exact_ties_on_cuts_walls_and_slabs(gpu):
    """Positions exactly on the cuts (owner by upper_bound - 1), on the box
    faces, just below the lower face (wraps to exactly the upper face, no
    fold-back, clamped to the last subdomain), and exactly on the slab
    thresholds lower + w / upper - w of a subdomain: migration and halo
    bit-exact with the oracle on a 2x3x2 grid (y non-periodic)."""
    torch = gpu
    from dsdecompx_amd.grid import SubdomainGrid
    from dsdecompx_amd.halo import BoxConfig
    say("grid ties")
    upper, periodic, w, dims = (8.0, 9.0, 6.0), (True, False, True), 0.5, (2, 3, 2)
    og = po.Grid(po.Box(upper=upper, periodic=periodic, width=w), dims)
    grid = SubdomainGrid(BoxConfig(upper=upper, periodic=periodic, width=w), dims, 4096)
    cuts = [og.cuts(a) for a in range(3)]
    vals = []
    for a in range(3):
        c = cuts[a]
        v = set(c) | {x + w for x in c} | {x - w for x in c}
        v |= {np.nextafter(x, -np.inf) for x in c} | {np.nextafter(x, np.inf) for x in c}
        if periodic[a]:
            v |= {-1e-17, -0.0}
        v = sorted(x for x in v if (periodic[a] or 0.0 <= x <= upper[a]))
        vals.append(v)
    rng = np.random.default_rng(2)
    pos = np.array([(rng.choice(vals[0]), rng.choice(vals[1]), rng.choice(vals[2])) for _ in range(1200)])
    n = len(pos) // og.ranks
    parts, owned = [], []
    for r in range(og.ranks):
        p = po.Particles(4096)
        chunk = pos[r * n:(r + 1) * n]
        p.x[:n], p.y[:n], p.z[:n] = chunk.T
        p.global_id[:n] = np.arange(r * n, (r + 1) * n)
        p.vx[:n] = rng.standard_normal(n)
        parts.append(p)
        owned.append(n)
    _load(torch, grid, parts, owned)
    og.migrate(parts, owned)
    assert grid.migrate() == owned
    _compare(grid, parts, owned, [0] * og.ranks, "ties migrate")
    ghosts = og.exchange_halo(parts, owned)
    assert grid.exchange_halo() == ghosts
    _compare(grid, parts, owned, ghosts, "ties halo")


def _random_case(seed):
    """A seeded grid configuration: dims (2..12 subdomains), a box with
    non-zero lower bounds and irrational-looking extents, a width up to a
    quarter of the smallest subdomain, random periodicity and neighbour
    policy."""
    rng = np.random.default_rng(seed)
    while True:
        dims = tuple(int(v) for v in rng.integers(1, 4, 3))
        if 2 <= dims[0] * dims[1] * dims[2] <= 12:
            break
    lower = tuple(float(v) for v in rng.uniform(-50.0, 50.0, 3))
    ext = rng.uniform(6.0, 14.0, 3) * np.array(dims)
    upper = tuple(float(l + e) for l, e in zip(lower, ext))
    w = float(rng.uniform(0.2, 0.25) * min(ext / np.array(dims)))
    periodic = tuple(bool(v) for v in rng.integers(0, 2, 3))
    if not any(periodic):
        periodic = (True, periodic[1], periodic[2])
    corners = bool(rng.integers(0, 2))
    return dims, lower, upper, periodic, w, corners


@pytest.mark.parametrize("seed", range(8))
def test_random_grids_with_shifted_boxes(gpu, seed):
    """Seeded grids on boxes that do not start at the origin: two steps of
    migrate / halo / reverse, every array bit-exact with the oracle."""
    torch = gpu
    from dsdecompx_amd.grid import SubdomainGrid
    from dsdecompx_amd.halo import BoxConfig
    dims, lower, upper, periodic, w, corners = _random_case(seed)
    say("random grid", seed, dims, periodic, corners)
    og = po.Grid(po.Box(lower=lower, upper=upper, periodic=periodic, width=w, edges_and_corners=corners), dims)
    n, cap = 800, 6000
    grid = SubdomainGrid(BoxConfig(lower=lower, upper=upper, periodic=periodic, width=w,
                                   edges_and_corners=corners), dims, cap)
    rng = np.random.default_rng(100 + seed)
    lo, hi = np.array(lower), np.array(upper)
    parts, owned = [], []
    for r in range(og.ranks):
        p = po.Particles(cap)
        pos = rng.uniform(lo - 0.1 * (hi - lo), hi + 0.1 * (hi - lo), (n, 3))
        for a in range(3):
            if not periodic[a]:
                pos[:, a] = np.clip(pos[:, a], lo[a], hi[a])
        p.x[:n], p.y[:n], p.z[:n] = pos.T
        for k in ("vx", "vy", "vz", "charge"):
            getattr(p, k)[:n] = rng.standard_normal(n)
        p.mass[:n] = 1.0
        p.type[:n] = rng.integers(0, 3, n)
        p.global_id[:n] = np.arange(r * n, (r + 1) * n)
        parts.append(p)
        owned.append(n)
    _load(torch, grid, parts, owned)
    for step in range(2):
        og.migrate(parts, owned)
        assert grid.migrate() == owned, f"seed {seed} step {step} owned"
        _compare(grid, parts, owned, [0] * og.ranks, f"seed {seed} step {step} migrate")
        ghosts = og.exchange_halo(parts, owned)
        assert grid.exchange_halo() == ghosts, f"seed {seed} step {step} ghosts"
        _compare(grid, parts, owned, ghosts, f"seed {seed} step {step} halo")
        f = rng.uniform(-1, 1, (og.ranks, 3, cap))
        for r, p in enumerate(parts):
            m = owned[r] + ghosts[r]
            for a, k in enumerate(("fx", "fy", "fz")):
                getattr(p, k)[:m] = f[r, a, :m]
                getattr(grid.particles, k)[grid.slots(r, m)] = torch.from_numpy(f[r, a, :m]).cuda()
        og.reverse_halo(parts, owned)
        grid.reverse_halo()
        _compare(grid, parts, owned, [0] * og.ranks, f"seed {seed} step {step} reverse", ("fx", "fy", "fz"))
        # move every owned particle a little for the next step
        for r, p in enumerate(parts):
            for a, k in enumerate(("x", "y", "z")):
                v = getattr(p, k)[:owned[r]]
                v += rng.normal(0, 0.4, owned[r])
                if not periodic[a]:
                    np.clip(v, lo[a], hi[a], out=v)
                getattr(grid.particles, k)[grid.slots(r, owned[r])] = torch.from_numpy(v.copy()).cuda()


def test_reference_run_on_eight_subdomains(gpu):
    """The 2x2x2 grid engine against the reference itself.  The fixture
    tests/golden/ref_grid_2x2x2.npz is what the reference's unmodified Host
    library held on each of 8 MPI ranks after every operation of two steps
    (tests/golden/make_ref_fixtures.py).  The box starts off 0 and its extents
    are not powers of two; particles start on the wrong ranks and past the
    faces; after the first migrate some sit exactly on slab thresholds.  The
    GPU must reproduce it bit for bit: owned ids, order, positions,
    velocities and a registered field after migrate; ghost ids, order and
    records after exchange_halo; owned forces after reverse_halo; the field
    after reverse_exchange(Sum)."""
    import os
    torch = gpu
    from dsdecompx_amd.grid import SubdomainGrid
    from dsdecompx_amd.halo import BoxConfig
    say("grid vs reference fixture")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_grid_2x2x2.npz")
    with np.load(path) as z:
        ref = {k: z[k] for k in z.files}
    box = BoxConfig(lower=tuple(ref["box_lower"]), upper=tuple(ref["box_upper"]),
                    width=float(ref["box_width"][0]))
    grid = SubdomainGrid(box, tuple(int(v) for v in ref["dims"]), 3000)
    grid.add_field("aux", torch.float64, 3)

    def put(name, r, values, count, offset=0):
        t = grid.field("aux") if name == "aux" else getattr(grid.particles, name)
        sl = grid.slots(r, offset + count)
        t[sl][offset:] = torch.from_numpy(np.ascontiguousarray(values)).cuda().reshape(t[sl][offset:].shape)

    owned = []
    for r in range(grid.ranks):
        n = len(ref[f"r{r}_s0_start_global_id"])
        for k in ("global_id", "type", "x", "y", "z", "vx", "vy", "vz", "mass", "charge"):
            put(k, r, ref[f"r{r}_s0_start_{k}"], n)
        put("aux", r, ref[f"r{r}_s0_start_aux"].reshape(n, 3), n)
        owned.append(n)
    grid.set_owned(owned)

    def check(name, r, want, count, offset=0, what=""):
        t = grid.field("aux") if name == "aux" else getattr(grid.particles, name)
        got = host(t[grid.slots(r, offset + count)][offset:]).ravel()
        assert_bits_equal(got, want.ravel(), what)

    for s in range(2):
        owned = grid.migrate()
        for r in range(grid.ranks):
            n, pre = owned[r], f"r{r}_s{s}_"
            assert n == len(ref[pre + "migrate_global_id"]), f"step {s} rank {r} owned count"
            for k in ("global_id", "x", "y", "z", "vx"):
                check(k, r, ref[pre + "migrate_" + k], n, what=f"step {s} rank {r} migrate {k}")
            check("aux", r, ref[pre + "migrate_aux"], n, what=f"step {s} rank {r} migrate aux")
            for k in ("x", "y", "z"):
                put(k, r, ref[pre + "ties_" + k], n)
        ghosts = grid.exchange_halo()
        for r in range(grid.ranks):
            n, g, pre = owned[r], ghosts[r], f"r{r}_s{s}_"
            assert g == len(ref[pre + "halo_global_id"]), f"step {s} rank {r} ghost count"
            for k in ("global_id", "x", "y", "z", "vz", "mass", "type"):
                check(k, r, ref[pre + "halo_" + k], g, n, f"step {s} rank {r} ghost {k}")
            check("aux", r, ref[pre + "halo_aux"], g, n, f"step {s} rank {r} ghost aux")
            for k in ("fx", "fy", "fz"):
                put(k, r, ref[pre + "forces_" + k], n + g)
        grid.reverse_halo()
        for r in range(grid.ranks):
            n, g, pre = owned[r], ghosts[r], f"r{r}_s{s}_"
            for k in ("fx", "fy", "fz"):
                check(k, r, ref[pre + "reverse_" + k], n, what=f"step {s} rank {r} reverse {k}")
            put("aux", r, ref[pre + "aux_ghosts"].reshape(g, 3), g, n)
        grid.reverse_exchange(["aux"])
        for r in range(grid.ranks):
            n, pre = owned[r], f"r{r}_s{s}_"
            check("aux", r, ref[pre + "aux_reduced"], n, what=f"step {s} rank {r} reduced aux")
            for k in ("x", "y", "z"):
                put(k, r, ref[pre + "displaced_" + k], n)
    grid.close()
