"""Batching: many replica boxes of uneven size in one launch per
phase.  Every replica's ghosts, positions and reverse sums are bit-exact with
the oracle run on that replica alone and with the plain per-box loop; the
full step captured in a hipGraph replays to the same bits as the eager step
(several replays, eager work in between)."""
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


def _specs(count, seed, nmin=2048, nmax=30000):
    from dsdecompx_amd.batch import ReplicaSpec
    from dsdecompx_amd.halo import BoxConfig
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(int)
    rho = rng.uniform(0.6, 1.2, count)
    specs = []
    for k in range(count):
        L = float((n[k] / rho[k]) ** (1.0 / 3.0))
        specs.append(ReplicaSpec(int(n[k]), BoxConfig(upper=(L, L, L), width=2.8)))
    return specs


def _fill(torch, b, seed):
    rng = np.random.default_rng(seed)
    for r, s in enumerate(b.specs):
        sl = b.owned_slice(r)
        L = s.box.upper[0]
        for k in ("x", "y", "z"):
            getattr(b, k)[sl] = torch.from_numpy(rng.uniform(0, L, s.owned)).cuda()
        for k in ("vx", "vy", "vz", "charge"):
            getattr(b, k)[sl] = torch.from_numpy(rng.standard_normal(s.owned)).cuda()
        b.mass[sl] = 1.0
        b.type[sl] = torch.from_numpy(rng.integers(0, 4, s.owned).astype(np.int32)).cuda()
        b.global_id[sl] = torch.arange(s.owned, dtype=torch.int32, device="cuda")


def _snapshot(b):
    return {k: host(getattr(b, k)).copy() for k in F64 + I32}


# 12 replicas: the owner-major passes walk the live tiles they count on the
# device; 80 replicas (more than a wave has lanes): the tiles numbered on the
# host, one per wave.  "long": one replica of 600,000 particles among small
# ones, in chunks of 256 with 256 scan threads, so that its count columns are
# scanned in 2 parts while the small replicas' second parts have nothing to do.
@pytest.mark.parametrize("count,seed,nmax,big,shape", [(12, 5, 30000, 0, None), (80, 31, 3000, 0, None),
                                                       (6, 7, 30000, 600000, (4, 4, 256))],
                         ids=["12", "80", "long"])
def test_batch_matches_oracle_and_loop(gpu, count, seed, nmax, big, shape):
    torch = gpu
    from dsdecompx_amd.batch import LoopedReplicas, ReplicaBatch, ReplicaSpec
    from dsdecompx_amd.halo import BoxConfig
    specs = _specs(count, seed, nmax=nmax)
    if big:
        L = float((big / 0.9) ** (1.0 / 3.0))
        specs[2] = ReplicaSpec(big, BoxConfig(upper=(L, L, L), width=2.8))
    say("batch", [s.owned for s in specs], shape)
    b = ReplicaBatch(specs)
    if shape:
        b.set_shape(*shape)
    _fill(torch, b, 6)
    start = _snapshot(b)
    rng = np.random.default_rng(9)
    disp = rng.normal(0, 0.1, (3, b.total))
    forces = rng.uniform(-1, 1, (3, b.total))

    # batched, eager
    for a, k in enumerate(("x", "y", "z")):
        getattr(b, k).add_(torch.from_numpy(disp[a]).cuda())
    b.migrate()
    b.exchange_halo()
    ghosts, owners, status = b.counts()
    assert not any(status)
    after_halo = _snapshot(b)
    own_all = host(b.ghost_owners())
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(b, k).copy_(torch.from_numpy(forces[a]).cuda())
    b.reverse_halo()
    after_rev = _snapshot(b)

    # oracle per replica
    for r, s in enumerate(specs):
        n, g, off = s.owned, ghosts[r], b.offsets[r]
        box = po.Box(upper=s.box.upper, width=s.box.width)
        o = po.Oracle(box)
        p = po.Particles(b.capacities[r])
        for k in F64 + I32:
            getattr(p, k)[:n] = start[k][off:off + n]
        for a, k in enumerate(("x", "y", "z")):
            getattr(p, k)[:n] += disp[a][off:off + n]
        o.migrate(p, n)
        assert o.exchange_halo(p, n) == g, f"replica {r} ghost count"
        for k in F64 + I32:
            assert_bits_equal(after_halo[k][off:off + n + g], getattr(p, k)[:n + g], f"r{r} halo {k}")
        assert_bits_equal(own_all[off:off + g], o.ghost_owners(), f"r{r} ghost owners")
        assert owners[r] == len(np.unique(o.ghost_owners()))
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(p, k)[:n + g] = forces[a][off:off + n + g]
        o.reverse_halo(p, n)
        for k in ("fx", "fy", "fz"):
            assert_bits_equal(after_rev[k][off:off + n], getattr(p, k)[:n], f"r{r} reverse {k}")

    # the plain per-box loop on the same storage
    for k in F64 + I32:
        getattr(b, k).copy_(torch.from_numpy(start[k]).cuda())
    loop = LoopedReplicas(b)
    for a, k in enumerate(("x", "y", "z")):
        getattr(b, k).add_(torch.from_numpy(disp[a]).cuda())
    loop.migrate()
    loop.exchange_halo()
    assert loop.ghosts == ghosts
    for r, s in enumerate(specs):
        off, n, g = b.offsets[r], s.owned, ghosts[r]
        for k in F64 + I32:
            assert_bits_equal(host(getattr(b, k)[off:off + n + g]), after_halo[k][off:off + n + g],
                              f"loop r{r} {k}")
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(b, k).copy_(torch.from_numpy(forces[a]).cuda())
    loop.reverse_halo()
    for r, s in enumerate(specs):
        off, n = b.offsets[r], s.owned
        for k in ("fx", "fy", "fz"):
            assert_bits_equal(host(getattr(b, k)[off:off + n]), after_rev[k][off:off + n],
                              f"loop r{r} reverse {k}")


def test_batch_512_replicas_full_size(gpu):
    """The batched workload at its configured size: 512 periodic replicas,
    particle counts log-uniform in [2048, 262144], density U[0.6, 1.2],
    w = 2.8, one N(0, 0.1) step of migrate -> exchange_halo -> reverse_halo
    (examples/communication_test/main.cu:222-275).  Every replica is checked
    bit for bit against the oracle run on that replica alone, and the plain
    per-box loop against the batch."""
    torch = gpu
    from dsdecompx_amd.batch import LoopedReplicas, ReplicaBatch
    specs = _specs(512, 2024, nmin=2048, nmax=262144)
    say("batch 512", sum(s.owned for s in specs), "particles")
    b = ReplicaBatch(specs)
    _fill(torch, b, 7)
    start = {k: host(getattr(b, k)).copy() for k in F64 + I32}
    rng = np.random.default_rng(8)
    disp = rng.normal(0, 0.1, (3, b.total))
    forces = rng.uniform(-1, 1, (3, b.total))
    for a, k in enumerate(("x", "y", "z")):
        getattr(b, k).add_(torch.from_numpy(disp[a]).cuda())
    b.migrate()
    b.exchange_halo()
    ghosts, owners, status = b.counts()
    assert not any(status)
    after_halo = _snapshot(b)
    own_all = host(b.ghost_owners())
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(b, k).copy_(torch.from_numpy(forces[a]).cuda())
    b.reverse_halo()
    after_rev = {k: host(getattr(b, k)).copy() for k in ("fx", "fy", "fz")}
    say("batch 512 oracle", sum(ghosts), "ghosts")
    for r, s in enumerate(specs):
        n, g, off = s.owned, ghosts[r], b.offsets[r]
        o = po.Oracle(po.Box(upper=s.box.upper, width=s.box.width))
        p = po.Particles(b.capacities[r])
        for k in F64 + I32:
            getattr(p, k)[:n] = start[k][off:off + n]
        for a, k in enumerate(("x", "y", "z")):
            getattr(p, k)[:n] += disp[a][off:off + n]
        o.migrate(p, n)
        assert o.exchange_halo(p, n) == g, f"replica {r} ghost count"
        for k in F64 + I32:
            assert_bits_equal(after_halo[k][off:off + n + g], getattr(p, k)[:n + g], f"r{r} halo {k}")
        assert_bits_equal(own_all[off:off + g], o.ghost_owners(), f"r{r} ghost owners")
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(p, k)[:n + g] = forces[a][off:off + n + g]
        o.reverse_halo(p, n)
        for k in ("fx", "fy", "fz"):
            assert_bits_equal(after_rev[k][off:off + n], getattr(p, k)[:n], f"r{r} reverse {k}")
    say("batch 512 loop")
    for k in F64 + I32:
        getattr(b, k).copy_(torch.from_numpy(start[k]).cuda())
    loop = LoopedReplicas(b)
    for a, k in enumerate(("x", "y", "z")):
        getattr(b, k).add_(torch.from_numpy(disp[a]).cuda())
    loop.migrate()
    loop.exchange_halo()
    assert loop.ghosts == ghosts
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(b, k).copy_(torch.from_numpy(forces[a]).cuda())
    loop.reverse_halo()
    got = _snapshot(b)
    for r, s in enumerate(specs):
        off, n, g = b.offsets[r], s.owned, ghosts[r]
        for k in ("x", "y", "z", "vx", "mass", "type", "global_id"):
            assert_bits_equal(got[k][off:off + n + g], after_halo[k][off:off + n + g], f"loop r{r} {k}")
        for k in ("fx", "fy", "fz"):
            assert_bits_equal(got[k][off:off + n], after_rev[k][off:off + n], f"loop r{r} reverse {k}")


@pytest.mark.parametrize("shape,fused", [((8, 4, 1024), False), ((4, 2, 256), True),
                                         ((16, 8, 512), False), ((8, 4, 1024), True)])
def test_graph_replays_match_eager(gpu, shape, fused):
    torch = gpu
    from dsdecompx_amd.batch import ReplicaBatch, capture_step
    specs = _specs(9, 11)
    say("graph", shape)
    b = ReplicaBatch(specs)
    b.set_shape(*shape)
    _fill(torch, b, 12)
    start = _snapshot(b)
    disp = [torch.randn((3, b.total), dtype=torch.float64, device="cuda",
                        generator=torch.Generator(device="cuda").manual_seed(40 + k)) * 0.1
            for k in range(3)]
    frc = torch.rand((3, b.total), dtype=torch.float64, device="cuda",
                     generator=torch.Generator(device="cuda").manual_seed(50)) * 2 - 1
    step_disp = torch.zeros((3, b.total), dtype=torch.float64, device="cuda")

    def app_halo():
        b.x.add_(step_disp[0]); b.y.add_(step_disp[1]); b.z.add_(step_disp[2])

    def app_rev():
        b.fx.copy_(frc[0]); b.fy.copy_(frc[1]); b.fz.copy_(frc[2])

    # eager reference: three steps
    eager = []
    for k in range(3):
        step_disp.copy_(disp[k])
        app_halo(); b.migrate(); b.exchange_halo(); app_rev(); b.reverse_halo()
        eager.append(_snapshot(b))
    # graph: restore, capture (warm-up steps run eagerly inside capture_step),
    # then restore again and replay
    for k_ in F64 + I32:
        getattr(b, k_).copy_(torch.from_numpy(start[k_]).cuda())
    step_disp.zero_()
    graph = capture_step(b, app_halo, app_rev, warmup=1, fused=fused)
    for k_ in F64 + I32:
        getattr(b, k_).copy_(torch.from_numpy(start[k_]).cuda())
    for k in range(3):
        step_disp.copy_(disp[k])
        graph.replay()
        torch.cuda.synchronize()
        got = _snapshot(b)
        for name in F64 + I32:
            assert_bits_equal(got[name], eager[k][name], f"replay {k} {name}")


def test_overflowing_replica_is_skipped(gpu):
    """A replica whose ghosts do not fit its storage gets its status word and
    is left untouched (ghost slots, forces); every other replica of the same
    launches is still bit-exact with the oracle."""
    torch = gpu
    from dsdecompx_amd.batch import ReplicaBatch
    specs = _specs(6, 21)
    bad = 2
    probe = ReplicaBatch(specs)
    caps = list(probe.capacities)
    probe.close()
    caps[bad] = specs[bad].owned + 8
    say("overflow replica", bad, caps[bad])
    b = ReplicaBatch(specs, capacities=caps)
    _fill(torch, b, 22)
    sentinel = float("nan")
    for r, s_ in enumerate(specs):
        sl = slice(b.offsets[r] + s_.owned, b.offsets[r] + b.capacities[r])
        for k in ("x", "y", "z", "vx", "fx"):
            getattr(b, k)[sl] = sentinel
    start = _snapshot(b)
    b.migrate()
    b.exchange_halo()
    ghosts, owners, status = b.counts()
    assert status[bad] != 0 and all(st == 0 for r, st in enumerate(status) if r != bad)
    rng = np.random.default_rng(23)
    forces = rng.uniform(-1, 1, (3, b.total))
    for r, s_ in enumerate(specs):
        # forces on owned slots everywhere; ghost slots of the good replicas
        n = s_.owned + (0 if r == bad else ghosts[r])
        sl = slice(b.offsets[r], b.offsets[r] + n)
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(b, k)[sl] = torch.from_numpy(forces[a][sl]).cuda()
    before_rev = _snapshot(b)
    b.reverse_halo()
    after = _snapshot(b)
    # the overflowing replica: ghost region and forces untouched
    lo, hi = b.offsets[bad], b.offsets[bad] + b.capacities[bad]
    for k in F64 + I32:
        assert_bits_equal(after[k][lo:hi], before_rev[k][lo:hi], f"overflowed replica {k}")
    n_bad = specs[bad].owned
    for k in ("x", "y", "z"):
        assert np.isnan(after[k][lo + n_bad:hi]).all()
    # the others: oracle
    for r, s_ in enumerate(specs):
        if r == bad:
            continue
        n, g, off = s_.owned, ghosts[r], b.offsets[r]
        o = po.Oracle(po.Box(upper=s_.box.upper, width=s_.box.width))
        p = po.Particles(b.capacities[r])
        for k in F64 + I32:
            getattr(p, k)[:n] = start[k][off:off + n]
        o.migrate(p, n)
        assert o.exchange_halo(p, n) == g
        for k in ("x", "y", "z", "vx", "vy", "vz", "mass", "charge", "type", "global_id"):
            assert_bits_equal(before_rev[k][off:off + n + g], getattr(p, k)[:n + g], f"r{r} {k}")
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(p, k)[:n + g] = forces[a][off:off + n + g]
        o.reverse_halo(p, n)
        for k in ("fx", "fy", "fz"):
            assert_bits_equal(after[k][off:off + n], getattr(p, k)[:n], f"r{r} reverse {k}")
