"""Process grids on one GPU (dsdx_grid_*): every subdomain's migration, halo
and reverse sums bit-exact with the oracle's Host path on the same number of
ranks (oracle/dsdx_oracle_grid.c), over several steps; and the reference
example getting_started on 2, 4 and 8 subdomains."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import F64, I32, assert_bits_equal, host, say

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
        f = rng.uniform(-1, 1, (og.ranks, 3, cap))
        for r, p in enumerate(parts):
            m = owned[r] + ghosts[r]
            for a, k in enumerate(("fx", "fy", "fz")):
                getattr(p, k)[:m] = f[r, a, :m]
                getattr(grid.particles, k)[grid.slots(r, m)] = torch.from_numpy(f[r, a, :m]).cuda()
        og.reverse_halo(parts, owned)
        grid.reverse_halo()
        _compare(grid, parts, owned, [0] * og.ranks, f"step {step} reverse", ("fx", "fy", "fz"))
