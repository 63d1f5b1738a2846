"""HIP graphs captured in C++ (dsdecompx_amd/bin/batch_bench, check mode):
the batched step over 512 replica boxes (particle counts 2,048-262,144),
captured once with hipStreamBeginCapture and replayed with hipGraphLaunch,
unfused and fused, and the eager batch and the fused per-box loop, all end
bit-identical to the plain per-box loop after three steps; and one box
captured right after dsdx_plan_reserve in global capture mode (any device
allocation or host wait inside the captured build would fail the capture)
replays bit-identical to the same steps run eagerly."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_captured_graphs_bit_exact():
    exe = os.path.join(ROOT, "dsdecompx_amd", "bin", "batch_bench")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run __graft_entry__.build() first")
    print("[case] batch_bench 512 replicas check", flush=True)
    r = subprocess.run([exe, "512", "3", "0", "2029", "check"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["bit_exact"]
    assert set(res["differing_words"]) == {"loop_fused", "batch", "graph", "batch_fused", "graph_fused",
                                           "single_box_graph_after_reserve"}
    assert all(v == 0 for v in res["differing_words"].values())
