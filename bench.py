#!/usr/bin/env python3
"""Benchmark of the loopback particle-communication path on one MI355X.

One step = the hot path migrate -> exchange_halo (rebuild) -> reverse_halo on
the synthetic workload; between the phases the application's own work
(a N(0, sigma) displacement before migrate, U[-1,1] forces on owned + ghosts
before reverse_halo) runs too but is not part of the measured hot path: HIP
events on the launch stream bracket migrate + exchange_halo and reverse_halo
separately, and their sum per step is `ms_per_step`.  The wall-clock time of
the whole loop (application work included) is reported as
`wall_ms_per_step`.  Per-kernel durations (`kernels`, `roofline`) come from
a second pass of the same number of steps in which every dispatch of the
library carries its own HIP start / stop events; those events delay each
dispatch by a few microseconds, so they are off in the timed pass.

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
    "uniform-16m": (16_777_216, 256.0, 2.8, 0.1, (True, True, True)),
    "uniform-16m-zwalls": (16_777_216, 256.0, 2.8, 0.1, (True, True, False)),
    "uniform-32k": (32_768, 32.0, 2.8, 0.05, (True, True, True)),
}


def model_bytes(n, g, u, periodic_axes):
    """The byte model the metric's roofline target is quoted on: migrate reads
    and writes x, y, z on every periodic axis (16 N per axis); the rebuild
    reads x, y, z (24 N) and, per ghost, the owner's 72-byte record and the
    96-byte ghost record (168 G); reverse_halo reads 24 B per ghost and reads
    + writes 48 B per owner.  It counts migrate's writes of every coordinate,
    which the path does not make (only changed coordinates are stored), so
    it is reported as the model's figure, never as a per-kernel fraction."""
    return {
        "migrate_wrap": 16.0 * n * periodic_axes,
        "halo_classify": 24.0 * n,
        "halo_scan": 0.0,
        "halo_scatter": 0.0,
        "halo_copy": 168.0 * g,
        "reverse_halo": 24.0 * g + 48.0 * u,
    }


def required_bytes(n, g, u, periodic_axes, wrapped):
    """Bytes each kernel must move: migrate reads the coordinates of every
    periodic axis (8 N per axis) and writes the `wrapped` coordinates that
    change (8 B each); the classification reads x, y, z (24 N); the ghost copy
    reads the owner's 72-byte record and writes the 96-byte ghost record per
    ghost (168 G); reverse_halo reads 24 B per ghost and reads + writes 48 B
    per boundary owner; the scan and route passes move plan bookkeeping only.
    The fused migrate + classification reads x, y, z once (24 N) and writes
    the changed coordinates."""
    return {
        "migrate_wrap": 8.0 * n * periodic_axes + 8.0 * wrapped,
        "halo_classify": 24.0 * n,
        "halo_scan": 0.0,
        "halo_scatter": 0.0,
        "halo_copy": 168.0 * g,
        "reverse_halo": 24.0 * g + 48.0 * u,
        "migrate_classify": 24.0 * n + 8.0 * wrapped,
    }


def step_required(bytes_, fused):
    """A step's required bytes: the split form's kernels, or the fused form's."""
    skip = ("migrate_wrap", "halo_classify") if fused else ("migrate_classify",)
    return sum(v for k, v in bytes_.items() if k not in skip)


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
            "ms_per_step": ms, "cpu_model": _cpu_model(),
            "sample": f"{name} workload ({n} particles), {timed} timed steps after 1 warm-up, "
                      f"oracle (C, -O2, one thread) on {os.cpu_count()} visible host cores"}


def cpu_reference(n, L, w, sigma, steps=1):
    """The reference's own Host path (its unmodified sources built by
    oracle/ref.mk into oracle/_ref/) on 8 MPI ranks: a 2x2x2 periodic grid of
    the same particle count and box, the reference's communication_test step
    (migrate, exchange_halo(rebuild), forces, reverse_halo), one warm-up and
    `steps` timed steps, the slowest rank's busy time.  The reference cannot
    run the single-box loopback itself (on one rank it has no ghosts), so its
    figure is for its own decomposition.  None when the build or MPICH is not
    here."""
    import json as _json
    import shutil
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_grid_fixture")
    mpiexec = "/opt/conda/bin/mpiexec" if os.path.exists("/opt/conda/bin/mpiexec") else shutil.which("mpiexec")
    if not os.path.exists(exe) or not mpiexec:
        return None
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = "/usr/lib/x86_64-linux-gnu:/opt/conda/lib:" + env.get("LD_LIBRARY_PATH", "")
    try:
        r = subprocess.run([mpiexec, "-n", "8", exe, "time", str(n), repr(L), repr(w), repr(sigma), str(steps)],
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env, cwd="/tmp",
                           timeout=240)
    except subprocess.TimeoutExpired:
        return {"error": "timed out after 240 s"}
    if r.returncode != 0:
        return {"error": f"exit {r.returncode}: {r.stderr.strip()[-300:]}"}
    res = _json.loads(r.stdout.strip().splitlines()[-1])
    ms = res["ms_per_step"]
    return {"value": n / (ms * 1e3), "unit": "Mparticles/s", "ms_per_step": ms, "cores": 8,
            "kind": "reference", "cpu_model": _cpu_model(),
            "sample": f"the reference's Host library (unmodified, oracle/ref.mk) on 8 MPI ranks, a 2x2x2 "
                      f"periodic grid of {n} particles in L={L}, w={w}, N(0,{sigma}) per step; "
                      f"{res['steps']} timed step(s) after 1 warm-up, slowest rank"}


def cpu_all_cores(name, n, L, w, sigma, steps, seed=11):
    """The oracle on every host core at once: P independent periodic
    sub-boxes of n / P particles each at the workload's density (the work of a
    P-rank Host run without its MPI transfers), one thread per sub-box, wall
    clock per step over all P.  A reported figure beside `cpu_baseline`."""
    import concurrent.futures as cf
    import numpy as np
    from oracle import pyoracle as po

    P = max(1, min(16, os.cpu_count() or 1))
    m = n // P
    Ls = L / P ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    work = []
    for _ in range(P):
        ora = po.Oracle(po.Box(upper=(Ls, Ls, Ls), width=w))
        cap = int(m * 1.5) + 4096
        p = po.Particles(cap)
        p.x[:m], p.y[:m], p.z[:m] = (rng.uniform(0, Ls, m) for _ in range(3))
        p.vx[:m], p.vy[:m], p.vz[:m] = (rng.standard_normal(m) for _ in range(3))
        p.mass[:m] = 1.0
        p.global_id[:m] = np.arange(m, dtype=np.int32)
        work.append((ora, p, rng.normal(0.0, sigma, (3, m)), rng.uniform(-1, 1, (3, cap))))

    ghosts = [0] * P

    def displace(r, k):
        ora, p, disp, force = work[r]
        for a, key in enumerate(("x", "y", "z")):
            getattr(p, key)[:m] += np.roll(disp[a], (k * 7919) % m)

    def build(r):
        ora, p, disp, force = work[r]
        ora.migrate(p, m)
        ghosts[r] = ora.exchange_halo(p, m)

    def fill(r):
        ora, p, disp, force = work[r]
        for a, key in enumerate(("fx", "fy", "fz")):
            getattr(p, key)[:m + ghosts[r]] = force[a, :m + ghosts[r]]

    def reverse(r):
        ora, p, disp, force = work[r]
        ora.reverse_halo(p, m)

    # Like cpu_baseline, the displacement and the force fill (numpy, under the
    # interpreter lock) are outside the timed phases.
    timed = 0.0
    with cf.ThreadPoolExecutor(P) as pool:
        for k in range(steps + 1):
            list(pool.map(lambda r: displace(r, k), range(P)))
            t0 = time.perf_counter()
            list(pool.map(build, range(P)))
            t1 = time.perf_counter()
            list(pool.map(fill, range(P)))
            t2 = time.perf_counter()
            list(pool.map(reverse, range(P)))
            t3 = time.perf_counter()
            if k > 0:
                timed += (t1 - t0) + (t3 - t2)
    ms = 1e3 * timed / steps
    return {"value": P * m / (ms * 1e3), "unit": "Mparticles/s", "cores": P, "kind": "port",
            "ms_per_step": ms,
            "sample": f"{name} workload split into {P} independent periodic boxes of {m} particles "
                      f"(L = {Ls:.1f}), one oracle thread each, {steps} timed steps after 1 warm-up"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measure_abi(args, n, L, w, sigma, periodic):
    """The hot path through the C ABI from Python (ctypes), torch tensors as
    device storage."""
    import torch
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig, DeviceParticles, LoopbackHalo

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
    # timed pass with the per-kernel timers off, then a pass with them on
    timers = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    ghosts, owners = [], []
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    for s in range(args.steps):
        ghosts.append(step(timers[s]))
        owners.append(_owners(halo))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    hot_ms = sum(t[0].elapsed_time(t[1]) + t[2].elapsed_time(t[3]) for t in timers)
    _abi.timing_reset()
    _abi.timing_enable(True)
    for s in range(args.steps):
        step()
    torch.cuda.synchronize()
    _abi.timing_enable(False)
    kern = _abi.timing_read()
    return {"hot_ms_per_step": hot_ms / args.steps, "wall_ms_per_step": 1e3 * wall / args.steps,
            "ghosts_per_step": sum(ghosts) / len(ghosts), "owners": sum(owners) / len(owners),
            "kernel_steps": args.steps,
            "kernels": {k: {"total_ms": v[0], "launches": v[1]} for k, v in kern.items()},
            "path": "C ABI from Python (dsdx_migrate_wrap, dsdx_halo_build, dsdx_reverse_halo)"}


def tune_shape(args, n, L, w, periodic):
    """Autotunes the launch shape on a synthetic box of the workload's size
    (dsdecompx_amd.autotune, HIP-event timed) and returns (shape, report)."""
    import torch
    from dsdecompx_amd import autotune
    from dsdecompx_amd.halo import BoxConfig, DeviceParticles, LoopbackHalo

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    torch.cuda.set_stream(torch.cuda.Stream(dev))
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed + 1)
    p = DeviceParticles(int(n * 1.12) + 4096, dev)
    for k in ("x", "y", "z"):
        getattr(p, k)[:n].uniform_(0.0, L, generator=gen)
    p.global_id[:n] = torch.arange(n, dtype=torch.int32, device=dev)
    halo = LoopbackHalo(BoxConfig(upper=(L, L, L), periodic=periodic, width=w))
    report = autotune.tune_single(halo, p, n)
    halo.close()
    del p
    torch.cuda.empty_cache()
    return report["best"], report


def run_comm_bench(args, n, L, w, sigma, periodic, env, mode="split"):
    import subprocess
    exe = os.path.join(ROOT, "dsdecompx_amd", "bin", "comm_bench")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`")
    cmd = [exe, str(n), repr(L), repr(w), repr(sigma), "1" if periodic[2] else "0",
           str(args.steps), str(args.warmup), str(args.seed), "0", str(args.steps), mode]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
    if r.returncode != 0:
        raise SystemExit(f"comm_bench failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def measure_cpp(args, n, L, w, sigma, periodic):
    """The hot path through the drop-in C++ API (dsdecompx::Decomp), driven
    by the compiled driver dsdecompx_amd/bin/comm_bench: migrate, then
    exchange_halo, then reverse_halo.  The same loop with
    Decomp::migrate_and_exchange_halo (the fused extension) runs after it
    and is reported under `fused`."""
    env = dict(os.environ)
    tuned = None
    if args.tune:
        from dsdecompx_amd import autotune
        shape, tuned = tune_shape(args, n, L, w, periodic)
        env.update(autotune.env_for(shape))
    res = run_comm_bench(args, n, L, w, sigma, periodic, env)
    res["env"] = env
    res["path"] = "C++ API (dsdecompx::Decomp migrate / exchange_halo / reverse_halo)"
    res["autotune"] = tuned
    if args.fused:
        res["fused_run"] = run_comm_bench(args, n, L, w, sigma, periodic, env, mode="fused")
    if args.refresh:
        res["refresh_runs"] = {m: run_comm_bench(args, n, L, w, sigma, periodic, env, mode=m)
                               for m in ("refresh", "refresh_pos")}
    if args.zwalls and periodic[2]:
        res["zwalls_run"] = run_comm_bench(args, n, L, w, sigma, (True, True, False), env)
    return res


def measure_fields(args):
    """The clustered workload with registered fields, through the drop-in C++
    API (comm_bench, mode "fields"):
    1,048,576 particles in 64 Gaussian clusters (sigma 4) in a periodic box
    L = 101.6, halo width 2.8, displacement N(0, 0.5) per step, and two
    registered fields (add_field) -- density (f64) and vec (3 x f64), F = 32 B
    per particle -- that travel into the ghosts with exchange_halo; per step
    migrate, exchange_halo, reverse_halo, reverse_exchange(Sum) of both
    fields.  Required bytes per step: migrate 24 N + 8 B per rewritten
    coordinate, halo 24 N + (168 + 2F) G, reverse 24 G + 48 U,
    reverse_exchange F G + 2F U."""
    n, L, w, sigma = 1048576, 101.6, 2.8, 0.5
    res = run_comm_bench(args, n, L, w, sigma, (True, True, True), dict(os.environ), mode="fields")
    ms, G, U = res["hot_ms_per_step"], res["ghosts_per_step"], res["owners"]
    C = res.get("wrapped_coords_per_step", 0.0)
    F = 32.0
    step_bytes = (24.0 * n + 8.0 * C + 24.0 * n + (168.0 + 2 * F) * G + 24.0 * G + 48.0 * U
                  + F * G + 2 * F * U)
    return {
        "workload": f"{n} particles in 64 Gaussian clusters (sigma 4) in a periodic box L={L}, "
                    f"halo width {w}, displacement N(0,{sigma}) per axis per step, 2 registered "
                    f"fields (f64 density, 3 x f64 vector), reverse_exchange(Sum) of both",
        "value": n / (ms * 1e3), "unit": "Mparticles/s", "ms_per_step": ms,
        "ghosts_per_step": G, "owners": U, "wrapped_coords_per_step": C,
        "step_required_bytes": step_bytes,
        "step_hbm_frac_of_peak": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "kernels": {k: {"avg_ms": v["total_ms"] / v["launches"], "launches": v["launches"]}
                    for k, v in res["kernels"].items() if v["launches"]},
    }


def measure_grid(args, n, L, w, sigma, dims=(2, 2, 2)):
    """The same particles split over a dims process grid, every subdomain on
    this GPU (dsdx_grid_*), driven by dsdecompx_amd/bin/grid_bench: migrate
    between subdomains, halo build, reverse_halo per step.  Algorithmic bytes
    per step: migrate 24 B (positions) + 192 B (every record is rewritten:
    retained particles are compacted, then arrivals appended) per particle;
    halo 24 B per particle + 168 B per ghost; reverse 24 B per ghost + 48 B
    per distinct owner."""
    import subprocess
    exe = os.path.join(ROOT, "dsdecompx_amd", "bin", "grid_bench")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`")
    steps = min(args.steps, 50)
    cmd = [exe, str(n), repr(L), repr(w), repr(sigma), *map(str, dims), str(steps),
           str(min(args.warmup, 5)), str(args.seed), "10"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise SystemExit(f"grid_bench failed ({r.returncode}): {r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    ms, G, U = res["hot_ms_per_step"], res["ghosts_per_step"], res["owners_per_step"]
    step_bytes = 216.0 * n + 24.0 * n + 168.0 * G + 24.0 * G + 48.0 * U
    return {
        "workload": f"{n} particles uniform in a periodic box L={L}, halo width {w}, displacement "
                    f"N(0,{sigma}); {'x'.join(map(str, dims))} process grid, every subdomain on this GPU",
        "dims": list(dims), "steps": steps,
        "value": n / (ms * 1e3), "unit": "Mparticles/s", "ms_per_step": ms,
        "ghosts_per_step": G, "owners_per_step": U,
        "owned_min": res["owned_min"], "owned_max": res["owned_max"],
        "step_algorithmic_bytes": step_bytes,
        "step_hbm_frac_of_peak": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "kernels": {k: {"avg_ms": v["total_ms"] / v["launches"], "launches": v["launches"]}
                    for k, v in res["kernels"].items() if v["launches"]},
    }


def measure_batched(args):
    """512 replica boxes, particle counts log-uniform in [2048, 262144],
    density U[0.6, 1.2], w = 2.8, displacement N(0, 0.1), through the C ABI
    from C++ (dsdecompx_amd/bin/batch_bench): the plain per-box loop, the
    batched form eager, and the batched step captured once with
    hipStreamBeginCapture and replayed with hipGraphLaunch, each unfused
    (migrate then exchange_halo) and fused (migrate_and_exchange_halo), with
    the same application work (displacement, forces) in every form.  A check
    run of the driver first confirms that every form ends bit-identical to the
    unfused loop."""
    import subprocess
    exe = os.path.join(ROOT, "dsdecompx_amd", "bin", "batch_bench")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`")
    R = args.replicas
    chk = subprocess.run([exe, str(R), "3", "0", str(args.seed + 5), "check"], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True)
    if chk.returncode not in (0, 1):
        raise SystemExit(f"batch_bench check failed ({chk.returncode}): {chk.stderr[-2000:]}")
    check = json.loads(chk.stdout.strip().splitlines()[-1])
    r = subprocess.run([exe, str(R), str(args.batch_steps), "2", str(args.seed + 5), "bench"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise SystemExit(f"batch_bench failed ({r.returncode}): {r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    ms, rs = res["ms_per_step"], res["replica_steps_per_s"]
    return {
        "workload": f"{R} periodic replica boxes, {res['particles']} particles (log-uniform "
                    f"2048..262144 per box, density U[0.6,1.2]), halo width 2.8, "
                    f"displacement N(0,0.1)",
        "replicas": R, "particles": res["particles"], "steps": res["steps"],
        "ms_per_step": ms, "replica_steps_per_s": rs,
        "graph_vs_looped": ms["loop"] / ms["graph"],
        "graph_vs_looped_fused": ms["loop_fused"] / ms["graph_fused"],
        "eager_vs_looped": ms["loop"] / ms["batch"],
        "graph_vs_eager": ms["batch"] / ms["graph"],
        "batch_shape": res.get("batch_shape"),
        "bit_exact_all_forms": check["bit_exact"],
        "check": check,
        "driver": "dsdecompx_amd/bin/batch_bench (C++, C ABI, hipStreamBeginCapture / hipGraphLaunch)",
        "timing": "wall clock over the steps, application displacement and force fill included in "
                  "every form; like for like: unfused against unfused, fused against fused",
    }


# Library kernel name -> device kernels (as rocprofv3 names them; Decomp::migrate
# launches wrap_box_kernel, the C ABI's per-array form migrate_wrap_kernel; the
# ghost copy and reverse of one box larger than the memory-side cache are
# copy_chunks_kernel and reverse_chunks_kernel, and its build runs no route
# pass (route_kernel runs in batches, grids and smaller boxes); a count column
# of up to 8 tiles is scanned by scan_parts_kernel, a longer one by
# scan_kernel).
DEVICE_KERNELS = {"migrate_wrap": ("wrap_box_kernel", "migrate_wrap_kernel"),
                  "halo_classify": ("classify_kernel",), "halo_scan": ("scan_parts_kernel", "scan_kernel"),
                  "halo_scatter": ("route_kernel",), "halo_copy": ("copy_chunks_kernel", "copy_runs_kernel", "copy_kernel"),
                  "reverse_halo": ("reverse_chunks_kernel", "reverse_runs_kernel", "reverse_forces_kernel")}
CALIBRATION_BYTES = 1 << 30


def _counter_means(path, counter=None):
    """Mean value per dispatch of `counter` (of the file's only counter when
    None), by plain kernel name."""
    import csv
    sums = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if counter is not None and row["Counter_Name"] != counter:
                continue
            name = row["Kernel_Name"].replace("void ", "").replace("(anonymous namespace)::", "")
            name = name.replace("dsdx::", "").split("(")[0].split("<")[0].strip()
            total, count = sums.get(name, (0.0, 0))
            sums[name] = (total + float(row["Counter_Value"]), count + 1)
    return {k: t / c for k, (t, c) in sums.items()}


def measure_traffic(n, L, w, sigma, periodic, env, seed):
    """HBM bytes per dispatch of every kernel of the step, from two rocprofv3
    counter runs of the C++ driver on the same workload (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads
    for comm_bench's plain copy of CALIBRATION_BYTES in the same run; and
    from three more runs the L2 (TCC) hit rate and the memory requests by
    size (reads are whole 128-byte lines on gfx950; writes go out as 64- or
    32-byte requests, the 32-byte ones being partly written lines).
    None when rocprofv3 is not on PATH."""
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    exe = os.path.join(ROOT, "dsdecompx_amd", "bin", "comm_bench")
    env = dict(env)
    env["TMPDIR"] = "/tmp"
    tmp = tempfile.mkdtemp(prefix="dsdx_pmc_", dir="/tmp")
    means, scale = {}, {}
    try:
        # one pass per group (the TCC block collects at most 4 counters at once)
        for group in (("FETCH_SIZE",), ("WRITE_SIZE",), ("TCC_HIT_sum", "TCC_MISS_sum"),
                      ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum"), ("TCC_EA0_RDREQ_sum",)):
            out_dir = os.path.join(tmp, group[0])
            cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", *group, "--output-format", "csv",
                   "-d", out_dir, "-o", "run", "--", exe, str(n), repr(L), repr(w), repr(sigma),
                   "1" if periodic[2] else "0", "5", "2", str(seed), str(CALIBRATION_BYTES), "0"]
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                               env=env, cwd="/tmp")
            files = glob.glob(os.path.join(out_dir, "**", "*counter_collection.csv"), recursive=True)
            if r.returncode != 0 or not files:
                if group[0] in ("FETCH_SIZE", "WRITE_SIZE"):
                    return {"error": f"rocprofv3 --pmc {group[0]} exited {r.returncode}: "
                                     f"{r.stderr.strip()[-300:]}"}
                continue
            for counter in group:
                means[counter] = _counter_means(files[0], counter)
            if group[0] in ("FETCH_SIZE", "WRITE_SIZE"):
                scale[group[0]] = means[group[0]]["calibration_copy"] * 1024.0 / CALIBRATION_BYTES
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    per_kernel = {}
    for lib_name, devs in DEVICE_KERNELS.items():
        for dev in devs:
            if dev not in means["FETCH_SIZE"] or dev not in means["WRITE_SIZE"]:
                continue
            rd = means["FETCH_SIZE"][dev] * 1024.0 / scale["FETCH_SIZE"]
            wr = means["WRITE_SIZE"][dev] * 1024.0 / scale["WRITE_SIZE"]
            entry = {"device_kernel": dev, "read_bytes": rd, "write_bytes": wr, "bytes": rd + wr}
            c = {k: means[k][dev] for k in ("TCC_HIT_sum", "TCC_MISS_sum", "TCC_EA0_WRREQ_sum",
                                             "TCC_EA0_WRREQ_64B_sum", "TCC_EA0_RDREQ_sum")
                 if k in means and dev in means[k]}
            if "TCC_HIT_sum" in c and "TCC_MISS_sum" in c and c["TCC_HIT_sum"] + c["TCC_MISS_sum"] > 0:
                entry["tcc_hit_rate"] = c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"])
            if "TCC_EA0_WRREQ_sum" in c and "TCC_EA0_WRREQ_64B_sum" in c:
                entry["ea_write_requests_64B"] = c["TCC_EA0_WRREQ_64B_sum"]
                entry["ea_write_requests_32B"] = c["TCC_EA0_WRREQ_sum"] - c["TCC_EA0_WRREQ_64B_sum"]
            if "TCC_EA0_RDREQ_sum" in c:
                entry["ea_read_requests"] = c["TCC_EA0_RDREQ_sum"]
            per_kernel[lib_name] = entry
            break
    return {"kernels": per_kernel,
            "scale": {"FETCH_SIZE": scale["FETCH_SIZE"], "WRITE_SIZE": scale["WRITE_SIZE"]},
            "method": f"rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, on comm_bench (5 steps after "
                      f"2 warm-up steps); KiB per dispatch x 1024 / (counter of a "
                      f"{CALIBRATION_BYTES}-byte plain copy x 1024 / {CALIBRATION_BYTES}); then "
                      f"TCC_HIT_sum + TCC_MISS_sum, TCC_EA0_WRREQ_sum + TCC_EA0_WRREQ_64B_sum and "
                      f"TCC_EA0_RDREQ_sum, one run each, per dispatch"}


def measure_wire(args, n, L, w):
    """The 16 M-particle halo rebuild and reverse through the communication records,
    as each rank of a transport-based multi-GPU run would execute them:
    dsdx_halo_build_records (classify, route, 96-byte ParticleRecords into the
    send buffer; no ghost written), the transfer (dsdx_loopback_exchange
    stands in for it), dsdx_unpack_records; then the returned force records:
    dsdx_pack_reverse_records, the transfer back, dsdx_accumulate_reverse_records.
    Against the direct path on the same particles: dsdx_halo_build (ghosts
    written in place, counts waited for) and dsdx_reverse_halo.  Buffers are
    allocated once; per-kernel averages come from the library timers.
    Algorithmic bytes: pack 48 B per owner (velocities, mass, charge, type,
    id; positions come densely from the classification pass) + 24 B per owner
    entry position + 96 B per record; loopback 2 x 96 B per record; unpack
    96 + 96 B per record; pack_reverse 24 + 96 B per record; accumulate 24 B
    per record + 48 B per owner."""
    import ctypes as C

    import torch
    from dsdecompx_amd import _abi
    from dsdecompx_amd.halo import BoxConfig, DeviceParticles, LoopbackHalo
    from dsdecompx_amd.wire import RECORD_BYTES, _buffer, _ptr

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.Stream(dev)
    torch.cuda.set_stream(stream)
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)
    cap = int(n * 1.15) + 4096
    p = DeviceParticles(cap, dev)
    for k in ("x", "y", "z"):
        getattr(p, k)[:n] = torch.rand(n, device=dev, dtype=torch.float64, generator=gen) * L
    for k in ("vx", "vy", "vz", "fx", "fy", "fz"):
        getattr(p, k)[:n] = torch.rand(n, device=dev, dtype=torch.float64, generator=gen) - 0.5
    p.mass[:n] = 1.0
    p.global_id[:n] = torch.arange(n, device=dev, dtype=torch.int32)
    halo = LoopbackHalo(BoxConfig(upper=(L, L, L), width=w))
    G = halo.exchange_halo(p, n)
    U = _owners(halo)
    lib = _abi.lib()
    s = stream.cuda_stream
    pc = p.c()
    box = halo._box
    send, recv = _buffer(G * RECORD_BYTES, dev), _buffer(G * RECORD_BYTES, dev)
    back, home = _buffer(G * RECORD_BYTES, dev), _buffer(G * RECORD_BYTES, dev)
    ptrs = [_ptr(b) for b in (send, recv, back, home)]
    fx, fy, fz = p.fx.data_ptr(), p.fy.data_ptr(), p.fz.data_ptr()
    count = C.c_uint64()

    def record_step():
        _abi.check(lib.dsdx_halo_build_records(halo.plan, C.byref(box), C.byref(pc), None, 0, n, ptrs[0], G,
                                               s, C.byref(count)))
        _abi.check(lib.dsdx_loopback_exchange(halo.plan, ptrs[0], ptrs[1], RECORD_BYTES, 0, s))
        _abi.check(lib.dsdx_unpack_records(ptrs[1], G, C.byref(pc), n, None, 0, s))
        _abi.check(lib.dsdx_pack_reverse_records(fx, fy, fz, n, G, ptrs[2], s))
        _abi.check(lib.dsdx_loopback_exchange(halo.plan, ptrs[2], ptrs[3], RECORD_BYTES, 1, s))
        _abi.check(lib.dsdx_accumulate_reverse_records(halo.plan, ptrs[3], fx, fy, fz, n, s))

    def direct_step():
        _abi.check(lib.dsdx_halo_build(halo.plan, C.byref(box), C.byref(pc), None, 0, n, cap, s,
                                       C.byref(count)))
        _abi.check(lib.dsdx_reverse_halo(halo.plan, fx, fy, fz, n, s))

    steps = max(10, min(args.steps, 100))
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        t0.record(stream)
        for _ in range(steps):
            fn()
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / steps

    record_ms = timed(record_step)
    direct_ms = timed(direct_step)
    assert count.value == G
    lib.dsdx_timing_reset()
    lib.dsdx_timing_enable(1)
    for _ in range(steps):
        record_step()
    torch.cuda.synchronize()
    lib.dsdx_timing_enable(0)
    algo = {"halo_pack": 72.0 * U + 96.0 * G, "loopback_exchange": 2 * 96.0 * G,
            "unpack_records": 192.0 * G, "pack_reverse": 120.0 * G,
            "accumulate_reverse": 24.0 * G + 48.0 * U, "halo_classify": 24.0 * n}
    kernels = {}
    for name in ("halo_classify", "halo_scan", "halo_scatter", "halo_pack", "loopback_exchange",
                 "unpack_records", "pack_reverse", "accumulate_reverse"):
        tot, cnt = C.c_double(), C.c_uint64()
        lib.dsdx_timing_read(name.encode(), C.byref(tot), C.byref(cnt))
        if cnt.value:
            avg = tot.value / cnt.value
            kernels[name] = {"avg_ms": avg, "launches": cnt.value}
            if name in algo:
                kernels[name]["algorithmic_bytes"] = algo[name]
                kernels[name]["algorithmic_GBps"] = algo[name] / (avg * 1e-3) / 1e9
    return {
        "workload": f"{n} particles uniform in a periodic box L={L}, halo width {w}; per step a halo "
                    f"rebuild and reverse_halo, through records (build_records, loopback, unpack, "
                    f"pack_reverse, loopback, accumulate) and directly (halo_build, reverse_halo)",
        "records": G, "owners": U, "steps": steps,
        "record_ms_per_step": record_ms, "direct_ms_per_step": direct_ms, "kernels": kernels,
    }


def measure_autograd(args):
    """65,536 particles, L = 40.3, w = 2.8, one rebuild: the autograd
    form -- owned positions through HaloPositions (positions-only refresh
    along the saved plan), a differentiable scalar sum(out * weight) of all
    positions, backward through the reverse_halo kernel -- against the plain
    calls that compute the same gradient (positions-only refresh, the weights
    written into the force arrays, reverse_halo).  Per-iteration times on
    the same stream, and whether the two gradients agree bit for bit."""
    import torch
    from dsdecompx_amd import autograd as ag
    from dsdecompx_amd.halo import BoxConfig, DeviceParticles, LoopbackHalo

    n, L, w = 65536, 40.3, 2.8
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    torch.cuda.set_stream(torch.cuda.Stream(dev))
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed + 9)
    p = DeviceParticles(2 * n + 4096, dev)
    for k in ("x", "y", "z"):
        getattr(p, k)[:n].uniform_(0.0, L, generator=gen)
    p.global_id[:n] = torch.arange(n, dtype=torch.int32, device=dev)
    halo = LoopbackHalo(BoxConfig(upper=(L, L, L), width=w))
    halo.migrate(p, n)
    g = ag.rebuild(halo, p, n)
    pos0 = torch.stack([p.x[:n], p.y[:n], p.z[:n]], 1).clone()
    weight = torch.randn((n + g, 3), dtype=torch.float64, device=dev, generator=gen)
    wcols = [weight[:, a].contiguous() for a in range(3)]

    def autograd_step():
        pos = pos0.clone().requires_grad_(True)
        out = ag.halo_positions(pos, halo, p)
        (out * weight).sum().backward()
        return pos.grad

    def plain_step():
        p.x[:n] = pos0[:, 0]
        p.y[:n] = pos0[:, 1]
        p.z[:n] = pos0[:, 2]
        halo.refresh(p, n, positions_only=True)
        p.fx[:n + g] = wcols[0]
        p.fy[:n + g] = wcols[1]
        p.fz[:n + g] = wcols[2]
        halo.reverse_halo(p, n)
        return torch.stack([p.fx[:n], p.fy[:n], p.fz[:n]], 1)

    def timed(fn, iters):
        for _ in range(5):
            fn()
        s0 = torch.cuda.Event(enable_timing=True)
        s1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        s0.record()
        for _ in range(iters):
            fn()
        s1.record()
        torch.cuda.synchronize()
        return s0.elapsed_time(s1) / iters

    iters = 100
    ms_auto = timed(autograd_step, iters)
    ms_plain = timed(plain_step, iters)
    same = torch.equal(autograd_step().view(torch.int64), plain_step().view(torch.int64))
    return {
        "workload": f"{n} particles uniform in L={L} (periodic), halo width {w}, {g} ghosts",
        "autograd_ms_per_iter": ms_auto, "plain_ms_per_iter": ms_plain,
        "plain_vs_autograd": ms_auto / ms_plain,
        "bit_exact_gradient": bool(same),
        "form": "autograd: HaloPositions forward (positions-only refresh) + sum(out*weight) + "
                "backward (reverse_halo kernel); plain: refresh + weights into forces + "
                "reverse_halo",
    }


def run(args):
    name = args.config
    n, L, w, sigma, periodic = WORKLOADS[name]
    if args.particles:
        n = args.particles
    if not (periodic[0] and periodic[1]):
        raise SystemExit("x and y are periodic in every workload")
    res = (measure_cpp if args.path == "cpp" else measure_abi)(args, n, L, w, sigma, periodic)
    ms = res["hot_ms_per_step"]
    G = res["ghosts_per_step"]
    U = res["owners"]
    C = res.get("wrapped_coords_per_step", 0.0)
    axes = sum(1 for a in periodic if a)
    mb = model_bytes(n, G, U, axes)
    step_bytes = sum(mb.values())
    ab = required_bytes(n, G, U, axes, C)
    step_req = step_required(ab, fused=False)

    per_kernel = {}
    for k, v in res["kernels"].items():
        if v["launches"]:
            per_kernel[k] = {"avg_ms": v["total_ms"] / v["launches"], "launches": v["launches"]}
            if ab.get(k):
                gbs = ab[k] / (per_kernel[k]["avg_ms"] * 1e-3) / 1e9
                per_kernel[k]["required_bytes"] = ab[k]
                per_kernel[k]["required_GBps"] = gbs
                per_kernel[k]["frac_of_peak"] = gbs / HBM_PEAK_GBS
    dominant = max((k for k in per_kernel if ab.get(k)), key=lambda k: per_kernel[k]["avg_ms"])
    dom_gbs = ab[dominant] / (per_kernel[dominant]["avg_ms"] * 1e-3) / 1e9

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
            "workload": f"{name}: {n} particles uniform in a box L={L} (periodic={list(periodic)}), "
                        f"halo width {w}, displacement N(0,{sigma}) per axis per step, "
                        f"1x1x1 self-loopback grid",
            "particles": n, "box": L, "halo_width": w, "periodic": list(periodic),
            "path": res["path"],
        },
        "roofline": {
            "bound": "hbm", "kernel": dominant,
            "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom_gbs / HBM_PEAK_GBS, "traffic": None,
            "algorithmic_bytes_per_launch": ab[dominant],
        },
        "step_hbm": {
            "required_bytes": step_req,
            "achieved_GBps": step_req / (ms * 1e-3) / 1e9,
            "frac_of_peak": step_req / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "basis": "bytes each kernel must move (required_bytes in bench.py)",
        },
        "step_hbm_model": {
            "model_bytes": step_bytes,
            "achieved_GBps": step_bytes / (ms * 1e-3) / 1e9,
            "frac_of_peak": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "target_ms_per_step_at_60pct": step_bytes / (0.6 * HBM_PEAK_GBS * 1e9) * 1e3,
            "basis": "the metric's byte model (migrate 16 N per periodic axis, halo 24 N + 168 G, "
                     "reverse 24 G + 48 U); it counts writes of unchanged coordinates",
        },
        "ghosts_per_step": G,
        "owners_per_step": U,
        "wrapped_coords_per_step": C,
        "wall_ms_per_step": res["wall_ms_per_step"],
        "kernel_pass_ms_per_step": res.get("kernel_pass_hot_ms_per_step"),
        "kernels": per_kernel,
        "autotune": res.get("autotune"),
    }
    if args.traffic and args.path == "cpp":
        t = measure_traffic(n, L, w, sigma, periodic, res.get("env", dict(os.environ)), args.seed)
        if t and "kernels" in t:
            for k, v in t["kernels"].items():
                if k in per_kernel:
                    per_kernel[k]["hbm_bytes"] = v["bytes"]
                    per_kernel[k]["hbm_read_bytes"] = v["read_bytes"]
                    per_kernel[k]["hbm_write_bytes"] = v["write_bytes"]
                    for extra in ("device_kernel", "tcc_hit_rate", "ea_read_requests",
                                  "ea_write_requests_64B", "ea_write_requests_32B"):
                        if extra in v:
                            per_kernel[k][extra] = v[extra]
                    if per_kernel[k].get("avg_ms"):
                        per_kernel[k]["hbm_GBps"] = v["bytes"] / (per_kernel[k]["avg_ms"] * 1e-3) / 1e9
            dom = t["kernels"].get(dominant)
            if dom:
                out["roofline"]["traffic"] = dom["bytes"]
                out["roofline"]["traffic_read"] = dom["read_bytes"]
                out["roofline"]["traffic_write"] = dom["write_bytes"]
            out["roofline"]["traffic_method"] = t["method"]
            out["roofline"]["traffic_scale"] = t["scale"]
        elif t:
            out["roofline"]["traffic_error"] = t["error"]
    fr = res.get("fused_run")
    if fr:
        fms = fr["hot_ms_per_step"]
        fab = required_bytes(n, fr["ghosts_per_step"], fr["owners"], axes,
                             fr.get("wrapped_coords_per_step", 0.0))
        fkern = {}
        for k, v in fr["kernels"].items():
            if v["launches"]:
                avg = v["total_ms"] / v["launches"]
                fkern[k] = {"avg_ms": avg, "launches": v["launches"]}
                if fab.get(k):
                    fkern[k]["required_GBps"] = fab[k] / (avg * 1e-3) / 1e9
                    fkern[k]["frac_of_peak"] = fkern[k]["required_GBps"] / HBM_PEAK_GBS
        out["fused"] = {
            "path": "C++ API extension Decomp::migrate_and_exchange_halo, then reverse_halo "
                    "(same results as the separate calls; one pass over the positions)",
            "value": n / (fms * 1e3), "unit": "Mparticles/s", "ms_per_step": fms,
            "step_hbm_frac_of_peak": step_required(fab, fused=True) / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "step_required_bytes": step_required(fab, fused=True),
            "step_hbm_model_frac_of_peak": step_bytes / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernels": fkern,
        }
    zr = res.get("zwalls_run")
    if zr:
        zms = zr["hot_ms_per_step"]
        zmb = model_bytes(n, zr["ghosts_per_step"], zr["owners"], 2)
        zab = required_bytes(n, zr["ghosts_per_step"], zr["owners"], 2,
                             zr.get("wrapped_coords_per_step", 0.0))
        out["z_walls"] = {
            "workload": f"{name} with z non-periodic (reflective walls at 0 and L applied by comm_bench)",
            "value": n / (zms * 1e3), "unit": "Mparticles/s", "ms_per_step": zms,
            "ghosts_per_step": zr["ghosts_per_step"], "owners_per_step": zr["owners"],
            "step_hbm_frac_of_peak": step_required(zab, fused=False) / (zms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "step_hbm_model_frac_of_peak": sum(zmb.values()) / (zms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        }
    rr = res.get("refresh_runs")
    if rr:
        # The reference benchmark's non-rebuild step: exchange_halo(rebuild = false) on the
        # plan of the last rebuild, then reverse_halo.  Bytes per refresh: full records
        # 168 B per ghost as in the rebuild's copy; positions only 24 B read (owner x, y, z)
        # + 48 B written (ghost x, y, z and zeroed forces) per ghost.
        out["refresh"] = {}
        for m, per_ghost in (("refresh", 168.0), ("refresh_pos", 72.0)):
            r = rr[m]
            k = r["kernels"].get("halo_refresh", {})
            avg = k["total_ms"] / k["launches"] if k.get("launches") else None
            g = r["last_ghosts"]
            out["refresh"]["full" if m == "refresh" else "positions_only"] = {
                "ms_per_step": r["hot_ms_per_step"],
                "value": n / (r["hot_ms_per_step"] * 1e3), "unit": "Mparticles/s",
                "refresh_kernel_avg_ms": avg, "ghosts": g,
                "refresh_algorithmic_GBps": per_ghost * g / (avg * 1e-3) / 1e9 if avg else None,
                "reverse_avg_ms": (r["kernels"]["reverse_halo"]["total_ms"] /
                                   r["kernels"]["reverse_halo"]["launches"]),
            }
    if args.fields:
        out["fields"] = measure_fields(args)
    if args.grid and all(periodic):
        out["grid"] = measure_grid(args, n, L, w, sigma)
    if args.wire and all(periodic):
        out["wire"] = measure_wire(args, n, L, w)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(name, n, L, w, sigma, periodic,
                                                                 args.cpu_steps)
    if out["cpu_baseline"] and all(periodic):
        out["cpu_baseline"]["reference_8_ranks"] = cpu_reference(n, L, w, sigma)
    if not args.no_cpu and all(periodic):
        out["cpu_all_cores"] = cpu_all_cores(name, n, L, w, sigma, args.cpu_steps)
    if not args.no_batched:
        out["batched"] = measure_batched(args)
    if args.autograd:
        out["autograd"] = measure_autograd(args)
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
    ap.add_argument("--config", default="uniform-16m", choices=sorted(WORKLOADS))
    ap.add_argument("--particles", type=int, default=0, help="override the particle count")
    ap.add_argument("--path", default="cpp", choices=("cpp", "abi"),
                    help="cpp: the C++ Decomp API driver; abi: the C ABI from Python")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched comparison (512 replica boxes)")
    ap.add_argument("--no-fused", dest="fused", action="store_false",
                    help="skip the fused migrate_and_exchange_halo run")
    ap.add_argument("--no-refresh", dest="refresh", action="store_false",
                    help="skip the reused-plan (refresh) runs")
    ap.add_argument("--no-autograd", dest="autograd", action="store_false",
                    help="skip the autograd-vs-plain comparison (65,536 particles)")
    ap.add_argument("--no-fields", dest="fields", action="store_false",
                    help="skip the run of 1,048,576 clustered particles with registered fields")
    ap.add_argument("--no-wire", dest="wire", action="store_false",
                    help="skip the communication-record (pack / unpack) run")
    ap.add_argument("--no-grid", dest="grid", action="store_false",
                    help="skip the 2x2x2 process-grid run of the same particles")
    ap.add_argument("--no-zwalls", dest="zwalls", action="store_false",
                    help="skip the z-non-periodic repeat of the workload")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the launch shape of the single box before measuring "
                         "(default: the library's default shape)")
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-traffic", dest="traffic", action="store_false",
                    help="skip the rocprofv3 counter runs that measure each kernel's HBM bytes "
                         "(use it when bench.py itself runs under a profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the loopback path has no "
                                   "multi-GPU transport; run with --gpus 1"}))
        sys.exit(2)
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
