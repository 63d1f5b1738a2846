"""The workloads of the benchmark at their real sizes, bit-exact against the
oracle on identical inputs:

- clustered, 1,048,576 particles in 64 Gaussian clusters (sigma 4, L = 101.6),
  displacement N(0, 0.5) per step (particles cross the box faces and wrap),
  two registered
  fields (density f64, vec3 f64) travelling with the ghosts and returned by
  reverse_exchange(Sum); three steps.
- uniform, 16,777,216 particles in L = 256, periodic and with z non-periodic
  (reflecting walls applied by the application); one full step.
- 3,000,000 particles (owned records larger than the 256 MB memory-side
  cache, so the ghost copy stages its runs through LDS): with an f64 and an
  f32 3-vector field, walking its units first to last and last to first, and
  with a halo so wide (w = 4 in L = 20) that most units hold more images than
  the staging map and are written one lane per owner; then the reused-plan
  refresh, positions only and the whole record with the fields.
- 2,500,000 and 5,000,000 particles in chunks of 256 with 256 scan threads:
  count columns scanned in 5 parts, and walked by one block per column.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import F64, I32, assert_bits_equal, compare_particles, host, say, synth, \
    to_device, to_oracle

from tests.gpu_util import box_ghost_writes  # noqa: F401  (autouse: every ghost-write kernel)
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


def test_clustered_fields_1m(gpu):
    torch = gpu
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    n, L, w = 1_048_576, 101.6, 2.8
    say("clustered 1M", n)
    d = synth(n, L, 2002, clusters=64, sigma=4.0)
    rng = np.random.default_rng(2003)
    cap = int(n * 1.6)
    dens = np.zeros(cap)
    dens[:n] = rng.uniform(0, 2, n)
    vec = np.zeros((cap, 3))
    vec[:n] = rng.uniform(-1, 1, (n, 3))
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
    op = to_oracle(d, cap)
    dev = to_device(d, cap)
    td = torch.from_numpy(dens.copy()).cuda()
    tv = torch.from_numpy(vec.copy()).cuda()
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    for step in range(3):
        say("clustered 1M step", step)
        disp = rng.normal(0, 0.5, (3, n))
        for a, k in enumerate(("x", "y", "z")):
            getattr(op, k)[:n] += disp[a]
            getattr(dev, k)[:n] += torch.from_numpy(disp[a]).cuda()
        crossed = int(((op.x[:n] < 0) | (op.x[:n] >= L) | (op.y[:n] < 0) | (op.y[:n] >= L) |
                       (op.z[:n] < 0) | (op.z[:n] >= L)).sum())
        assert crossed > 0, crossed   # particles leave the box and are wrapped back
        ora.migrate(op, n)
        halo.migrate(dev, n)
        g = ora.exchange_halo(op, n, fields=[po.Field(dens), po.Field(vec)])
        assert halo.exchange_halo(dev, n, fields=[td, tv]) == g
        compare_particles(dev, op, n + g, f"clustered step {step}")
        assert_bits_equal(host(td[:n + g]), dens[:n + g], "density")
        assert_bits_equal(host(tv[:n + g]).reshape(-1), vec[:n + g].reshape(-1), "vec")
        contrib = rng.uniform(0, 1, g)
        vc = rng.uniform(-1, 1, (g, 3))
        dens[n:n + g] = contrib
        vec[n:n + g] = vc
        td[n:n + g] = torch.from_numpy(contrib).cuda()
        tv[n:n + g] = torch.from_numpy(vc).cuda()
        ora.reverse_exchange(n, [po.Field(dens, po.SUM), po.Field(vec, po.SUM)])
        halo.reverse_exchange([(td, _abi.SUM), (tv, _abi.SUM)], n)
        assert_bits_equal(host(td[:n]), dens[:n], "density reverse")
        assert_bits_equal(host(tv[:n]).reshape(-1), vec[:n].reshape(-1), "vec reverse")
        f = rng.uniform(-1, 1, (3, n + g))
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(op, k)[:n + g] = f[a]
            getattr(dev, k)[:n + g] = torch.from_numpy(f[a]).cuda()
        ora.reverse_halo(op, n)
        halo.reverse_halo(dev, n)
        for k in ("fx", "fy", "fz"):
            assert_bits_equal(host(getattr(dev, k)[:n]), getattr(op, k)[:n], f"clustered reverse {k}")


@pytest.mark.parametrize("zper", [True, False], ids=["periodic", "z_walls"])
def test_uniform_16m_full_size_step(gpu, zper):
    torch = gpu
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    n, L, w = 16_777_216, 256.0, 2.8
    per = (True, True, zper)
    say("uniform 16M full size", per)
    rng = np.random.default_rng(303 if zper else 304)
    cap = int(n * 1.1)
    d = {k: rng.uniform(0, L, n) for k in ("x", "y", "z")}
    for k in ("vx", "vy", "vz", "charge", "fx", "fy", "fz"):
        d[k] = rng.standard_normal(n)
    d["mass"] = np.ones(n)
    d["type"] = np.zeros(n, np.int32)
    d["global_id"] = np.arange(n, dtype=np.int32)
    disp = rng.normal(0, 0.1, (3, n))
    for a, k in enumerate(("x", "y", "z")):
        d[k] = d[k] + disp[a]
    if not zper:   # reflecting walls, applied by the application
        z = d["z"]
        z[z < 0] = -z[z < 0]
        z[z > L] = 2 * L - z[z > L]
    op = to_oracle(d, cap)
    dev = to_device(d, cap)
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w, periodic=per))
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w, periodic=per))
    ora.migrate(op, n)
    halo.migrate(dev, n)
    g = ora.exchange_halo(op, n)
    assert halo.exchange_halo(dev, n) == g
    expect = ((L + 2 * w) ** 2 * (L + (2 * w if zper else 0)) - L ** 3) / L ** 3 * n
    assert abs(g - expect) < 0.01 * expect, (g, expect)
    compare_particles(dev, op, n + g, "uniform 16M halo")
    f = rng.uniform(-1, 1, (3, n + g))
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(op, k)[:n + g] = f[a]
        getattr(dev, k)[:n + g] = torch.from_numpy(f[a]).cuda()
    ora.reverse_halo(op, n)
    halo.reverse_halo(dev, n)
    for k in ("fx", "fy", "fz"):
        assert_bits_equal(host(getattr(dev, k)[:n]), getattr(op, k)[:n], f"uniform 16M reverse {k}")


@pytest.mark.parametrize("L,w,descending", [(160.0, 2.8, 0), (160.0, 2.8, 8), (20.0, 4.0, 0)],
                         ids=["runs", "runs_last_to_first", "wide_halo_per_owner"])
def test_large_box_ghost_runs(gpu, monkeypatch, L, w, descending):
    torch = gpu
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    n = 3_000_000
    monkeypatch.setenv("DSDX_DESCENDING", str(1 | 16 | descending))
    say("large box ghost runs", n, L, w, descending)
    d = synth(n, L, 4040 + int(L) + descending)
    rng = np.random.default_rng(4041)
    images = ((1 + 2 * w / L) ** 3 - 1) * n
    cap = int(n + 1.1 * images) + 4096
    dens = np.zeros(cap)
    dens[:n] = rng.uniform(0, 2, n)
    vec = np.zeros((cap, 3), np.float32)
    vec[:n] = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
    op = to_oracle(d, cap)
    dev = to_device(d, cap)
    td = torch.from_numpy(dens.copy()).cuda()
    tv = torch.from_numpy(vec.copy()).cuda()
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    for step in range(2):
        say("large box step", step)
        disp = rng.normal(0, 0.1, (3, n))
        for a, k in enumerate(("x", "y", "z")):
            getattr(op, k)[:n] += disp[a]
            getattr(dev, k)[:n] += torch.from_numpy(disp[a]).cuda()
        ora.migrate(op, n)
        halo.migrate(dev, n)
        g = ora.exchange_halo(op, n, fields=[po.Field(dens), po.Field(vec)])
        assert halo.exchange_halo(dev, n, fields=[td, tv]) == g
        assert abs(g - images) < 0.02 * images, (g, images)
        compare_particles(dev, op, n + g, f"large box step {step}")
        assert_bits_equal(host(td[:n + g]), dens[:n + g], "density")
        assert_bits_equal(host(tv[:n + g]).reshape(-1).view(np.int32), vec[:n + g].reshape(-1).view(np.int32),
                          "vec f32")
        f = rng.uniform(-1, 1, (3, n + g))
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(op, k)[:n + g] = f[a]
            getattr(dev, k)[:n + g] = torch.from_numpy(f[a]).cuda()
        ora.reverse_halo(op, n)
        halo.reverse_halo(dev, n)
        for k in ("fx", "fy", "fz"):
            assert_bits_equal(host(getattr(dev, k)[:n]), getattr(op, k)[:n], f"large box reverse {k}")
    # the last plan again: owners move a little, the ghosts re-read them
    for full in (False, True):
        say("large box refresh full", full)
        dx = rng.normal(0, 0.05, (3, n))
        for a, k in enumerate(("x", "y", "z")):
            getattr(op, k)[:n] += dx[a]
            getattr(dev, k)[:n] = torch.from_numpy(getattr(op, k)[:n].copy()).cuda()
        if full:
            op.vx[:n] += 1.0
            dev.vx[:n] = torch.from_numpy(op.vx[:n].copy()).cuda()
            dens[:n] += 1.0
            td[:n] = torch.from_numpy(dens[:n].copy()).cuda()
        before = {k: host(getattr(dev, k)[:n + g]).copy() for k in F64 + I32}
        dens_ref, vec_ref = host(td).copy(), host(tv).copy()
        halo.refresh(dev, n, fields=[td, tv] if full else (), positions_only=not full)
        ref = to_oracle({k: getattr(op, k)[:n].copy() for k in F64 + I32}, n + g)
        for k in F64 + I32:
            getattr(ref, k)[n:n + g] = before[k][n:]
        ora.write_ghosts(ref, n, fields=[po.Field(dens_ref), po.Field(vec_ref)] if full else ())
        if not full:   # positions_only: only x, y, z and zero forces change
            for k in ("vx", "vy", "vz", "mass", "charge", "type", "global_id"):
                getattr(ref, k)[n:n + g] = before[k][n:]
        compare_particles(dev, ref, n + g, f"large box refresh full={full}")
        assert_bits_equal(host(td[:n + g]), dens_ref[:n + g], "refresh density")
        assert_bits_equal(host(tv[:n + g]), vec_ref[:n + g], "refresh vec f32")


@pytest.mark.parametrize("n", [2_500_000, 5_000_000], ids=["five_parts", "column_walk"])
def test_long_count_columns(gpu, monkeypatch, n):
    """Chunks of 256 particles and 256 scan threads: a part covers 2,048
    chunks, so 2.5 M particles scan each count column in 5 parts, and 5 M
    particles (more than 8 parts) take the one-block-per-column scan.  The
    build and reverse_halo are bit-exact with the oracle either way."""
    torch = gpu
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    monkeypatch.setenv("DSDX_ITEMS", "4")
    monkeypatch.setenv("DSDX_SCAN_THREADS", "256")
    L, w = float(round(n ** (1.0 / 3.0), 1)), 2.8
    say("long count columns", n, L)
    d = synth(n, L, 6060 + n % 97)
    rng = np.random.default_rng(6061)
    cap = int(n * 1.25) + 4096
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
    op = to_oracle(d, cap)
    dev = to_device(d, cap)
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    g = ora.exchange_halo(op, n)
    assert halo.exchange_halo(dev, n) == g
    compare_particles(dev, op, n + g, f"long columns n={n}")
    f = rng.uniform(-1, 1, (3, n + g))
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(op, k)[:n + g] = f[a]
        getattr(dev, k)[:n + g] = torch.from_numpy(f[a]).cuda()
    ora.reverse_halo(op, n)
    halo.reverse_halo(dev, n)
    for k in ("fx", "fy", "fz"):
        assert_bits_equal(host(getattr(dev, k)[:n]), getattr(op, k)[:n], f"long columns reverse {k}")
