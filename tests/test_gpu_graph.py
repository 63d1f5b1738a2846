"""HIP graphs captured in C++ (dsdecompx_amd/bin/batch_bench, check mode):
the batched step over 512 replica boxes (particle counts 2,048-262,144),
captured once with hipStreamBeginCapture and replayed with hipGraphLaunch,
unfused and fused, and the eager batch and the fused per-box loop, all end
bit-identical to the plain per-box loop after three steps; and one box
captured right after dsdx_plan_reserve in global capture mode (any device
allocation or host wait inside the captured build would fail the capture)
replays bit-identical to the same steps run eagerly.  A captured single-box
build that asks for its ghost count fails at once instead of waiting."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("runs", ["default", "staged_runs"])
def test_cpp_captured_graphs_bit_exact(runs):
    """With staged_runs (DSDX_GHOST_RUNS=1) the single box's captured build
    takes the chunk units (no route pass) and the batch the tile units."""
    exe = os.path.join(ROOT, "dsdecompx_amd", "bin", "batch_bench")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run __graft_entry__.build() first")
    env = dict(os.environ)
    if runs == "staged_runs":
        env["DSDX_GHOST_RUNS"] = "1"
    print(f"[case] batch_bench 512 replicas check, {runs}", flush=True)
    r = subprocess.run([exe, "512", "3", "0", "2029", "check"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300, env=env)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["bit_exact"]
    assert set(res["differing_words"]) == {"loop_fused", "batch", "graph", "batch_fused", "graph_fused",
                                           "single_box_graph_after_reserve"}
    assert all(v == 0 for v in res["differing_words"].values())


def test_captured_build_cannot_wait_for_its_counts():
    """A single-box build captured into a graph cannot report its ghost
    count: asked for it, the call returns DSDX_ESTATE at once (nothing waits
    on counts that a capture never produces), the capture still ends cleanly,
    and the next eager build of the same plan waits for its own counts and
    matches the oracle bit for bit."""
    import ctypes as C

    import torch
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    from oracle import pyoracle as po
    from tests.gpu_util import compare_particles, say, synth, to_device, to_oracle

    n, L, w = 20000, 27.0, 2.8
    say("captured build with a count output", n)
    d = synth(n, L, 77)
    dev = to_device(d, 3 * n)
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    s = torch.cuda.Stream()
    first = halo.exchange_halo(dev, n, stream=s)   # eager: sizes the plan
    s.synchronize()
    graph = C.c_void_p()
    _abi.check(_abi.lib().dsdx_graph_begin(C.c_void_p(s.cuda_stream)), "graph begin")
    try:
        with pytest.raises(RuntimeError, match="captured"):
            halo.exchange_halo(dev, n, stream=s)
    finally:
        rc = _abi.lib().dsdx_graph_end(C.c_void_p(s.cuda_stream), C.byref(graph))
    _abi.check(rc, "graph end")
    _abi.check(_abi.lib().dsdx_graph_destroy(graph), "graph destroy")
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
    op = to_oracle(d, 3 * n)
    g = ora.exchange_halo(op, n)
    assert halo.exchange_halo(dev, n, stream=s) == g == first
    s.synchronize()
    compare_particles(dev, op, n + g, "eager build after a captured one")


def test_count_wait_behind_a_long_kernel():
    """The host waits for a build's counts by polling the stamped words the
    scan stores; after 50 ms it also asks the stream, every 10 ms, whether it
    failed or drained.  A build queued behind a few hundred ms of other work on
    the same stream takes that path and still returns the oracle's ghost
    count and ghosts."""
    import time

    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    from oracle import pyoracle as po
    from tests.gpu_util import compare_particles, say, synth, to_device, to_oracle

    n, L, w = 20000, 27.0, 2.8
    say("count wait behind a long kernel", n)
    d = synth(n, L, 78)
    dev = to_device(d, 3 * n)
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    s = torch.cuda.Stream()
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
    op = to_oracle(d, 3 * n)
    g = ora.exchange_halo(op, n)
    # a spin kernel of a few hundred ms on this GPU (calibrated roughly)
    with torch.cuda.stream(s):
        t0 = time.perf_counter()
        torch.cuda._sleep(10_000_000)
        s.synchronize()
        per_cycle = (time.perf_counter() - t0) / 10_000_000
        torch.cuda._sleep(int(0.4 / per_cycle))
    t0 = time.perf_counter()
    assert halo.exchange_halo(dev, n, stream=s) == g
    waited = time.perf_counter() - t0
    s.synchronize()
    say("waited", round(waited * 1e3, 1), "ms")
    assert waited > 0.05, "the build did not wait behind the spin kernel"
    compare_particles(dev, op, n + g, "build behind a long kernel")


def test_replayed_route_free_build_then_owner_maps(monkeypatch):
    """A single-box build over chunk units captured into a graph (staged
    runs forced), replayed after the positions moved, then eager calls that
    walk the owner maps (ghost owners, a full refresh): the graph routes with
    the build, so the maps are those of the replayed build and every result
    matches the oracle on the moved positions."""
    import ctypes as C

    import numpy as np
    import torch
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    from oracle import pyoracle as po
    from tests.gpu_util import F64, I32, assert_bits_equal, compare_particles, host, say, synth, to_device, \
        to_oracle

    monkeypatch.setenv("DSDX_GHOST_RUNS", "1")
    n, L, w = 30000, 30.0, 2.8
    say("replayed route-free build", n)
    d = synth(n, L, 91)
    dev = to_device(d, 3 * n)
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    s = torch.cuda.Stream()
    halo.exchange_halo(dev, n, stream=s)   # eager: sizes the plan
    s.synchronize()
    graph = C.c_void_p()
    _abi.check(_abi.lib().dsdx_graph_begin(C.c_void_p(s.cuda_stream)), "graph begin")
    try:
        halo.exchange_halo(dev, n, stream=s, wait=False)
    finally:
        rc = _abi.lib().dsdx_graph_end(C.c_void_p(s.cuda_stream), C.byref(graph))
    _abi.check(rc, "graph end")
    try:
        rng = np.random.default_rng(5)
        for step in range(3):
            # move the owned particles inside the box, then replay the build
            for k in ("x", "y", "z"):
                v = np.clip(d[k] + rng.normal(0.0, 0.3, n), 0.0, np.nextafter(L, 0.0))
                d[k] = v
                getattr(dev, k)[:n] = torch.from_numpy(v).cuda()
            torch.cuda.synchronize()
            _abi.check(_abi.lib().dsdx_graph_launch(graph, C.c_void_p(s.cuda_stream)), "graph launch")
            s.synchronize()
            ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
            op = to_oracle(d, 3 * n)
            g = ora.exchange_halo(op, n)
            compare_particles(dev, op, n + g, f"replay {step}")
            with torch.cuda.stream(s):
                own = host(halo.ghost_owners(g, stream=s))
            assert_bits_equal(own, ora.ghost_owners(), f"ghost owners after replay {step}")
            for k in ("x", "vx"):
                getattr(dev, k)[:n] += 0.001
                getattr(op, k)[:n] += 0.001
            torch.cuda.synchronize()
            halo.refresh(dev, n, stream=s)
            s.synchronize()
            ref = to_oracle({k: getattr(op, k)[:n].copy() for k in F64 + I32}, n + g)
            ora.write_ghosts(ref, n)
            compare_particles(dev, ref, n + g, f"refresh after replay {step}")
            for k in ("x", "vx"):
                d[k] = getattr(op, k)[:n].copy()
    finally:
        _abi.check(_abi.lib().dsdx_graph_destroy(graph), "graph destroy")
