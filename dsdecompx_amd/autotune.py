"""Launch-shape autotuner of the halo kernels.

A shape is (items per lane, waves per block, scan threads): a chunk is
64 x items consecutive particles handled by one wave; blocks hold `waves`
chunks; the per-replica count scan runs `scan threads` per column.  The
tuner times the rebuild (classify + scan + scatter) and reverse_halo of every
candidate on the caller's own data with HIP events, and keeps the fastest.
The rebuild is idempotent (it rewrites the same ghosts) and reverse_halo is
timed on scratch force arrays, so the caller's data is unchanged.

Results depend only on the particle counts and box sizes, so a tuned shape
can be reused for later steps; ``DSDX_ITEMS``, ``DSDX_WAVES`` and
``DSDX_SCAN_THREADS`` hand it to other processes (e.g. the C++ driver).
"""
from __future__ import annotations

import ctypes as C
import itertools

import torch

from . import _abi

ITEMS = (4, 8, 16)
WAVES = (2, 4, 8)
SCAN_THREADS = (256, 1024)
CANDIDATES = tuple(itertools.product(ITEMS, WAVES, SCAN_THREADS))


def _time(fn, reps: int) -> float:
    stream = torch.cuda.current_stream()
    fn()   # warm-up (and first-touch of resized buffers)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        fn()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def tune_batch(batch, reps: int = 5, candidates=CANDIDATES) -> dict:
    """Times every candidate shape on a ReplicaBatch; leaves the fastest set.
    Returns {"best": shape, "ms": {shape: ms}}."""
    scratch = [torch.zeros_like(batch.fx) for _ in range(3)]
    stream = torch.cuda.current_stream().cuda_stream
    results = {}

    def step():
        batch.exchange_halo()
        _abi.check(_abi.lib().dsdx_batch_reverse_halo(batch.handle, scratch[0].data_ptr(),
                                                      scratch[1].data_ptr(),
                                                      scratch[2].data_ptr(), stream))

    for shape in candidates:
        batch.set_shape(*shape)
        results[shape] = _time(step, reps)
    best = min(results, key=results.get)
    batch.set_shape(*best)
    batch.exchange_halo()
    torch.cuda.synchronize()
    return {"best": best, "ms": {f"{i}x{w}/{s}": v for (i, w, s), v in results.items()}}


def tune_single(halo, particles, owned: int, reps: int = 5, candidates=CANDIDATES) -> dict:
    """Same for one LoopbackHalo plan over DeviceParticles."""
    scratch = [torch.zeros_like(particles.fx) for _ in range(3)]
    stream = torch.cuda.current_stream().cuda_stream
    results = {}

    def step():
        halo.exchange_halo(particles, owned, wait=False)
        _abi.check(_abi.lib().dsdx_reverse_halo(halo.plan, scratch[0].data_ptr(),
                                                scratch[1].data_ptr(), scratch[2].data_ptr(),
                                                owned, stream))

    halo.exchange_halo(particles, owned)   # sizes the storage once, outside the timing
    for shape in candidates:
        _abi.check(_abi.lib().dsdx_plan_set_shape(halo.plan, *shape))
        results[shape] = _time(step, reps)
    best = min(results, key=results.get)
    _abi.check(_abi.lib().dsdx_plan_set_shape(halo.plan, *best))
    halo.exchange_halo(particles, owned)
    return {"best": best, "ms": {f"{i}x{w}/{s}": v for (i, w, s), v in results.items()}}


def env_for(shape) -> dict:
    items, waves, scan = shape
    return {"DSDX_ITEMS": str(items), "DSDX_WAVES": str(waves), "DSDX_SCAN_THREADS": str(scan)}


def current_shape(halo) -> tuple:
    i, w, s = C.c_int32(), C.c_int32(), C.c_int32()
    _abi.check(_abi.lib().dsdx_plan_get_shape(halo.plan, C.byref(i), C.byref(w), C.byref(s)))
    return i.value, w.value, s.value
