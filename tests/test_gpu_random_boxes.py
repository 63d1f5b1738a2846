"""Seeded random single boxes, each one step of migrate -> exchange_halo ->
reverse_halo and a reused-plan refresh through the C ABI, bit-exact with the
oracle: particle counts log-uniform from 1 to 3 M, density 0.5-1.5,
boxes away from the origin, random periodicity and neighbour policy, widths
up to more than half the box, displacements that carry particles across the
faces, and a random launch shape (items per lane, waves, scan threads) and
walk order per box -- so the count columns are scanned in one part, in
several and by the column walk, depending on the box."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import F64, I32, assert_bits_equal, compare_particles, host, say, synth, \
    to_device, to_oracle

from tests.gpu_util import box_ghost_writes  # noqa: F401  (autouse: every ghost-write kernel)
pytestmark = pytest.mark.gpu

SEEDS = list(range(12))


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


def _case(seed):
    rng = np.random.default_rng(9000 + seed)
    n = int(np.exp(rng.uniform(0.0, np.log(3e6))))
    rho = rng.uniform(0.5, 1.5)
    L = float(np.round(max(2.0, (n / rho) ** (1.0 / 3.0)), 2))
    w = float(np.round(rng.uniform(0.3, min(6.0, 0.7 * L)), 3))
    lower = float(np.round(rng.uniform(-40.0, 40.0), 1))
    periodic = tuple(bool(rng.uniform() < 0.75) for _ in range(3))
    edges = bool(rng.uniform() < 0.8)
    sigma = float(rng.uniform(0.0, 0.5))
    shape = (int(rng.choice([4, 8, 16])), int(rng.choice([2, 4, 8])), int(rng.choice([256, 1024])),
             int(rng.integers(0, 128)))
    return n, L, w, lower, periodic, edges, sigma, shape


@pytest.mark.parametrize("seed", SEEDS)
def test_random_box_step(gpu, monkeypatch, seed):
    torch = gpu
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    n, L, w, lower, periodic, edges, sigma, shape = _case(seed)
    say("random box", seed, n, L, w, lower, periodic, edges, sigma, shape)
    items, waves, scan, descending = shape
    monkeypatch.setenv("DSDX_ITEMS", str(items))
    monkeypatch.setenv("DSDX_WAVES", str(waves))
    monkeypatch.setenv("DSDX_SCAN_THREADS", str(scan))
    monkeypatch.setenv("DSDX_DESCENDING", str(descending))
    upper = tuple(lower + L for _ in range(3))
    box_o = po.Box(lower=(lower,) * 3, upper=upper, periodic=periodic, width=w, edges_and_corners=edges)
    box_d = BoxConfig(lower=(lower,) * 3, upper=upper, periodic=periodic, width=w, edges_and_corners=edges)
    d = synth(n, L, 9100 + seed, lower=lower)
    rng = np.random.default_rng(9200 + seed)
    for k in ("x", "y", "z"):
        d[k] = d[k] + rng.normal(0.0, sigma, n)
    # the oracle first, on ample storage, to size the device storage exactly
    ora = po.Oracle(box_o)
    op = to_oracle(d, n * 27 + 64)
    ora.migrate(op, n)
    g = ora.exchange_halo(op, n)
    dev = to_device(d, n + g + 64)
    halo = LoopbackHalo(box_d)
    halo.migrate(dev, n)
    assert halo.exchange_halo(dev, n) == g
    compare_particles(dev, op, n + g, f"random box {seed} halo")
    f = rng.uniform(-1.0, 1.0, (3, n + g))
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(op, k)[:n + g] = f[a]
        getattr(dev, k)[:n + g] = torch.from_numpy(f[a]).cuda()
    ora.reverse_halo(op, n)
    halo.reverse_halo(dev, n)
    for k in ("fx", "fy", "fz"):
        assert_bits_equal(host(getattr(dev, k)[:n]), getattr(op, k)[:n], f"random box {seed} reverse {k}")
    # the owners move a little; a full refresh rewrites every ghost from its owner
    for k in ("x", "y", "z"):
        getattr(op, k)[:n] += rng.normal(0.0, 0.05, n)
        getattr(dev, k)[:n] = torch.from_numpy(getattr(op, k)[:n].copy()).cuda()
    op.vx[:n] += 1.0
    dev.vx[:n] = torch.from_numpy(op.vx[:n].copy()).cuda()
    halo.refresh(dev, n, positions_only=False)
    ora.write_ghosts(op, n)
    compare_particles(dev, op, n + g, f"random box {seed} refresh")
