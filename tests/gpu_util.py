"""Shared helpers of the GPU parity tests: seeded synthetic particles, host <->
device copies, and the oracle run on the same inputs."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import pyoracle as po

F64 = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "charge")
I32 = ("type", "global_id")


def say(*args):
    print("[case]", *args, flush=True)
    sys.stdout.flush()


def synth(n, L, seed, lower=0.0, clusters=0, sigma=4.0):
    """Seeded owned particles: positions uniform (or Gaussian clusters,
    wrapped into the box), velocities N(0,1), mass 1, charge N(0,1),
    type 0..3, global_id 0..n-1."""
    rng = np.random.default_rng(seed)
    if clusters:
        centres = rng.uniform(lower, lower + L, (clusters, 3))
        pick = rng.integers(0, clusters, n)
        pos = centres[pick] + rng.normal(0.0, sigma, (n, 3))
        pos = lower + np.mod(pos - lower, L)
        pos = np.where(pos >= lower + L, lower, pos)
    else:
        pos = rng.uniform(lower, lower + L, (n, 3))
    d = {"x": pos[:, 0].copy(), "y": pos[:, 1].copy(), "z": pos[:, 2].copy()}
    for k in ("vx", "vy", "vz", "charge"):
        d[k] = rng.standard_normal(n)
    for k in ("fx", "fy", "fz"):
        d[k] = rng.uniform(-1, 1, n)
    d["mass"] = np.ones(n)
    d["type"] = rng.integers(0, 4, n).astype(np.int32)
    d["global_id"] = np.arange(n, dtype=np.int32)
    return d


def to_oracle(d, capacity):
    n = len(d["x"])
    p = po.Particles(capacity)
    for k in F64 + I32:
        getattr(p, k)[:n] = d[k]
    return p


def to_device(d, capacity):
    import torch
    from dsdecompx_amd.halo import DeviceParticles
    n = len(d["x"])
    p = DeviceParticles(capacity)
    for k in F64 + I32:
        getattr(p, k)[:n] = torch.from_numpy(np.ascontiguousarray(d[k])).to(p.device)
    return p


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def assert_bits_equal(a, b, what):
    """Bit-exact comparison (NaN payloads and -0.0 included)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} != {b.shape}"
    if a.dtype == np.float64:
        av, bv = a.view(np.uint64), b.view(np.uint64)
    elif a.dtype == np.float32:
        av, bv = a.view(np.uint32), b.view(np.uint32)
    else:
        av, bv = a, b
    bad = np.nonzero(av != bv)[0]
    assert bad.size == 0, (f"{what}: {bad.size} mismatches, first at {bad[0]}: "
                           f"{a[bad[0]]!r} vs {b[bad[0]]!r}")


def compare_particles(dev, ora, count, what):
    for k in F64 + I32:
        assert_bits_equal(host(getattr(dev, k)[:count]), getattr(ora, k)[:count], f"{what}.{k}")


@pytest.fixture(autouse=True, params=["per_owner", "runs"])
def ghost_writes(request, monkeypatch):
    """Runs a test module's tests twice: the ghost copy, refresh and
    reverse_halo working one lane per owner (DSDX_GHOST_RUNS=0) and staging
    each unit's slot runs through LDS (DSDX_GHOST_RUNS=1; reverse_halo follows
    it), whatever the sizes would pick.  The variable is read when a plan,
    batch or grid is created."""
    monkeypatch.setenv("DSDX_GHOST_RUNS", "1" if request.param == "runs" else "0")
    return request.param


@pytest.fixture(autouse=True, params=["per_owner", "runs", "tile_runs"])
def box_ghost_writes(request, monkeypatch):
    """The single-box modules' form of ``ghost_writes``, with a third case.
    With the staged runs a single box's build copies, and reverse_halo
    reduces, over units of whole chunks with no route pass ("runs", the
    default); "tile_runs" (DSDX_UNIT_OWNERS=0) takes the units of owner tiles
    after a route pass, as batches and process grids do."""
    monkeypatch.setenv("DSDX_GHOST_RUNS", "0" if request.param == "per_owner" else "1")
    if request.param == "tile_runs":
        monkeypatch.setenv("DSDX_UNIT_OWNERS", "0")
    return request.param
