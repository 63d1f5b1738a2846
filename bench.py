#!/usr/bin/env python3
"""Benchmark of the loopback particle-communication path on one MI355X.

One step = the hot path migrate -> exchange_halo (rebuild) -> reverse_halo on
the synthetic workload; between the phases the application's own work
(a N(0, sigma) displacement before migrate, U[-1,1] forces on owned + ghosts
before reverse_halo) runs too but is not part of the measured hot path: HIP
events on the launch stream bracket migrate + exchange_halo and reverse_halo
separately, and their sum per step is `ms_per_step`.  The wall-clock time of
the whole loop (application work included) is reported as
`wall_ms_per_step`.

Prints ONE JSON line (rank 0).  See --help for the workloads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "migrate+halo+reverse Mparticles/s at 16M FP64; % HBM peak; batched vs looped"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (particles, box length, halo width, displacement sigma, periodic)
    "cfg3": (16_777_216, 256.0, 2.8, 0.1, (True, True, True)),
    "cfg3z": (16_777_216, 256.0, 2.8, 0.1, (True, True, False)),
    "cfg1": (32_768, 32.0, 2.8, 0.05, (True, True, True)),
}


def algorithmic_bytes(n, g, u, periodic_axes):
    """Bytes each kernel must move by the path's definition: migrate reads
    and writes x,y,z on every periodic axis (16 N per axis); the rebuild
    reads x,y,z (24 N) in the classify pass and, per ghost, reads the owner's
    72-byte record and writes the 96-byte ghost record in the scatter pass;
    reverse_halo reads 24 B per ghost and reads + writes 48 B per owner."""
    return {
        "migrate_wrap": 16.0 * n * periodic_axes,
        "halo_classify": 24.0 * n,
        "halo_scan": 0.0,
        "halo_scatter": 168.0 * g,
        "reverse_halo": 24.0 * g + 48.0 * u,
    }


def cpu_baseline(name, n, L, w, sigma, periodic, steps, seed=7):
    """The oracle (single-threaded C restatement of the Host path) timed on
    this host over `steps` steps of the same workload."""
    import numpy as np
    from oracle import pyoracle as po

    rng = np.random.default_rng(seed)
    box = po.Box(upper=(L, L, L), periodic=periodic, width=w)
    ora = po.Oracle(box)
    cap = int(n * 1.12) + 4096
    p = po.Particles(cap)
    p.x[:n], p.y[:n], p.z[:n] = (rng.uniform(0, L, n) for _ in range(3))
    p.vx[:n], p.vy[:n], p.vz[:n] = (rng.standard_normal(n) for _ in range(3))
    p.mass[:n] = 1.0
    p.global_id[:n] = np.arange(n, dtype=np.int32)
    disp = rng.normal(0.0, sigma, (3, n))
    force = rng.uniform(-1, 1, (3, cap))
    total = 0.0
    timed = 0
    for step in range(steps + 1):
        shift = (step * 7919) % n
        for a, k in enumerate(("x", "y", "z")):
            getattr(p, k)[:n] += np.roll(disp[a], shift)
        t0 = time.perf_counter()
        ora.migrate(p, n)
        g = ora.exchange_halo(p, n)
        t1 = time.perf_counter()
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(p, k)[:n + g] = force[a, :n + g]
        t2 = time.perf_counter()
        ora.reverse_halo(p, n)
        t3 = time.perf_counter()
        if step > 0:
            total += (t1 - t0) + (t3 - t2)
            timed += 1
    ms = 1e3 * total / timed
    return {"value": n / (ms * 1e3), "unit": "Mparticles/s", "cores": 1, "kind": "port",
            "ms_per_step": ms,
            "sample": f"{name} workload ({n} particles), {timed} timed steps after 1 warm-up, "
                      f"oracle (C, -O2, one thread) on {os.cpu_count()} visible host cores"}


def run(args):
    import torch
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig, DeviceParticles, LoopbackHalo

    name = args.config
    n, L, w, sigma, periodic = WORKLOADS[name]
    if args.particles:
        n = args.particles
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    torch.cuda.set_stream(torch.cuda.Stream(dev))   # the path requires a non-null stream
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)

    cap = int(n * 1.12) + 4096
    p = DeviceParticles(cap, dev)
    for k in ("x", "y", "z"):
        getattr(p, k)[:n].uniform_(0.0, L, generator=gen)
    for k in ("vx", "vy", "vz"):
        getattr(p, k)[:n].normal_(0.0, 1.0, generator=gen)
    p.mass[:n] = 1.0
    p.global_id[:n] = torch.arange(n, dtype=torch.int32, device=dev)
    halo = LoopbackHalo(BoxConfig(upper=(L, L, L), periodic=periodic, width=w))
    disp = torch.empty((3, n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    def step(timers=None):
        disp.normal_(0.0, sigma, generator=gen)
        p.x[:n] += disp[0]
        p.y[:n] += disp[1]
        p.z[:n] += disp[2]
        if timers:
            timers[0].record(stream)
        halo.migrate(p, n)
        g = halo.exchange_halo(p, n)
        if timers:
            timers[1].record(stream)
        for k in ("fx", "fy", "fz"):
            getattr(p, k)[:n + g].uniform_(-1.0, 1.0, generator=gen)
        if timers:
            timers[2].record(stream)
        halo.reverse_halo(p, n)
        if timers:
            timers[3].record(stream)
        return g

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    _abi.timing_reset()
    _abi.timing_enable(True)
    timers = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    ghosts, owners = [], []
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    for s in range(args.steps):
        g = step(timers[s])
        ghosts.append(g)
        owners.append(_owners(halo))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    _abi.timing_enable(False)
    kern = _abi.timing_read()

    hot_ms = sum(t[0].elapsed_time(t[1]) + t[2].elapsed_time(t[3]) for t in timers)
    ms = hot_ms / args.steps
    G = sum(ghosts) / len(ghosts)
    U = sum(owners) / len(owners)
    axes = sum(1 for a in periodic if a)
    ab = algorithmic_bytes(n, G, U, axes)
    step_bytes = sum(ab.values())

    per_kernel = {}
    for k, (tot, cnt) in kern.items():
        if cnt:
            per_kernel[k] = {"avg_ms": tot / cnt, "launches": cnt}
    dominant = max((k for k in per_kernel if k in ab), key=lambda k: per_kernel[k]["avg_ms"])
    dom_ms = per_kernel[dominant]["avg_ms"]
    dom_gbs = ab[dominant] / (dom_ms * 1e-3) / 1e9
    for k in per_kernel:
        if k in ab and ab[k] > 0:
            per_kernel[k]["algorithmic_GBps"] = ab[k] / (per_kernel[k]["avg_ms"] * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": n / (ms * 1e3),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{name}: {n} particles uniform in a periodic box L={L} "
                        f"(periodic={list(periodic)}), halo width {w}, displacement N(0,{sigma}) "
                        f"per axis per step, 1x1x1 self-loopback grid",
            "particles": n, "box": L, "halo_width": w, "periodic": list(periodic),
            "path": "C ABI (dsdx_migrate_wrap, dsdx_halo_build, dsdx_reverse_halo)",
        },
        "roofline": {
            "bound": "hbm", "kernel": dominant,
            "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom_gbs / HBM_PEAK_GBS, "traffic": args.traffic,
        },
        "step_hbm": {
            "algorithmic_bytes": step_bytes,
            "achieved_GBps": step_bytes / (ms * 1e-3) / 1e9,
            "frac_of_peak": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "ghosts_per_step": G,
        "owners_per_step": U,
        "wall_ms_per_step": 1e3 * wall / args.steps,
        "kernels": per_kernel,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(name, n, L, w, sigma, periodic, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    return out


def _owners(halo):
    import ctypes as C
    from dsdecompx_amd import _abi
    u = C.c_uint64()
    _abi.check(_abi.lib().dsdx_plan_query(halo.plan, None, None, None, C.byref(u), None, None))
    return u.value


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="cfg3", choices=sorted(WORKLOADS))
    ap.add_argument("--particles", type=int, default=0, help="override the particle count")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the dominant kernel from rocprofv3 counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the loopback path has no "
                                   "multi-GPU transport; run with --gpus 1"}))
        sys.exit(2)
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
