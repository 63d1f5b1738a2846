"""GPU parity of the single-box copy and reverse_halo over chunk units (no
route pass in the build) at extreme unit sizes, against the CPU oracle:
one chunk per unit, about 64 owners per unit, and units of up to 64 chunks
whose owners or images overflow the LDS stage (written and reduced one lane
per owner).  Each build is followed by calls that need the owner maps, which
the route pass then writes on demand: the ghost-owner query, a refresh."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import assert_bits_equal, compare_particles, host, say
from tests.test_gpu_parity import CASES, _run_build

pytestmark = pytest.mark.gpu

UNIT_CASES = [c for c in CASES if c[0] in ("ragged_chunk", "uniform_32k", "z_nonperiodic", "wide_halo",
                                            "dense_small", "shifted_box")]


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dsdecompx_amd import _abi
    _abi.lib()
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


@pytest.fixture(params=[1, 64, 1024], ids=["unit1", "unit64", "unit1024"])
def unit_owners(request, monkeypatch):
    monkeypatch.setenv("DSDX_GHOST_RUNS", "1")
    monkeypatch.setenv("DSDX_UNIT_OWNERS", str(request.param))
    return request.param


@pytest.mark.parametrize("case", UNIT_CASES, ids=[c[0] for c in UNIT_CASES])
def test_build_reverse_then_owner_maps(gpu, unit_owners, case):
    import torch
    say("chunk units", case, "unit owners", unit_owners)
    ora, op, g_ref, halo, dev, g, n = _run_build(case, seed=11)
    assert g == g_ref
    compare_particles(dev, op, n + g, case[0])
    # reverse_halo first: it needs no owner maps
    rng = np.random.default_rng(17)
    for k in ("fx", "fy", "fz"):
        vals = rng.uniform(-1, 1, n + g)
        getattr(op, k)[:n + g] = vals
        getattr(dev, k)[:n + g] = torch.from_numpy(vals).to(dev.device)
    ora.reverse_halo(op, n)
    halo.reverse_halo(dev, n)
    for k in ("fx", "fy", "fz"):
        assert_bits_equal(host(getattr(dev, k)[:n + g]), getattr(op, k)[:n + g], k)
    # then the owner maps, routed on demand
    assert_bits_equal(host(halo.ghost_owners(g)), ora.ghost_owners(), "ghost owners")


@pytest.mark.parametrize("case", UNIT_CASES, ids=[c[0] for c in UNIT_CASES])
def test_refresh_after_route_free_build(gpu, unit_owners, case):
    """A full refresh right after a build over chunk units: the refresh is
    the first call that walks the owner maps.  The oracle rewrites the ghosts
    of its stored plan from the moved owners."""
    import torch
    from tests.gpu_util import F64, I32, to_oracle
    say("chunk units refresh", case, "unit owners", unit_owners)
    ora, op, g_ref, halo, dev, g, n = _run_build(case, seed=12)
    rng = np.random.default_rng(19)
    for k in ("x", "y", "z", "vx"):
        getattr(op, k)[:n] += rng.normal(0.0, 0.01, n)
        getattr(dev, k)[:n] = torch.from_numpy(getattr(op, k)[:n].copy()).to(dev.device)
    halo.refresh(dev, n, positions_only=False)
    ref = to_oracle({k: getattr(op, k)[:n].copy() for k in F64 + I32}, n + g)
    ora.write_ghosts(ref, n)
    compare_particles(dev, ref, n + g, case[0] + " refresh")


def test_rebuild_changes_unit_size(gpu, monkeypatch):
    """Two builds of one plan: the first sizes its units before any count is
    known, the second from the first's counts; then a build of another
    density on the same plan."""
    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    from tests.gpu_util import synth, to_device, to_oracle
    monkeypatch.setenv("DSDX_GHOST_RUNS", "1")
    say("chunk units rebuild")
    L, w, n = 40.0, 2.8, 60000
    box_o = po.Box(lower=(0.0,) * 3, upper=(L,) * 3, periodic=(True,) * 3, width=w, edges_and_corners=True)
    halo = LoopbackHalo(BoxConfig(lower=(0.0,) * 3, upper=(L,) * 3, periodic=(True,) * 3, width=w,
                                  edges_and_corners=True))
    for seed, spread in ((21, L), (22, L), (23, 0.25 * L)):
        d = synth(n, spread, seed)
        ora = po.Oracle(box_o)
        op = to_oracle(d, n * 3)
        g_ref = ora.exchange_halo(op, n)
        dev = to_device(d, n * 3)
        g = halo.exchange_halo(dev, n)
        assert g == g_ref
        compare_particles(dev, op, n + g, f"seed {seed}")
        rng = np.random.default_rng(seed)
        for k in ("fx", "fy", "fz"):
            vals = rng.uniform(-1, 1, n + g)
            getattr(op, k)[:n + g] = vals
            getattr(dev, k)[:n + g] = torch.from_numpy(vals).to(dev.device)
        ora.reverse_halo(op, n)
        halo.reverse_halo(dev, n)
        for k in ("fx", "fy", "fz"):
            assert_bits_equal(host(getattr(dev, k)[:n + g]), getattr(op, k)[:n + g], f"{k} seed {seed}")
