"""Writes the golden fixtures of tests/golden/*.npz from the pinned CPU oracle.

Each fixture is one seeded step of the path on a small box: the inputs
(positions before the displacement, the displacement, every particle array,
forces for owned + ghosts, registered fields) and the expected arrays after
migrate, after exchange_halo (ghosts appended) and after reverse_halo /
reverse_exchange.  tests/test_golden.py checks that the oracle still
reproduces them; tests/test_gpu_golden.py checks the HIP path against them.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from tests.gpu_util import F64, I32, synth, to_oracle  # noqa: E402

CASES = {
    # name: (n, L, width, periodic, edges, lower, clusters, sigma_move, fields)
    "uniform_periodic": (2048, 16.0, 2.8, (True, True, True), True, 0.0, 0, 0.05, False),
    "clusters_fields_znp": (1536, 20.0, 2.8, (True, True, False), True, 0.0, 4, 0.5, True),
    "faces_only_shifted": (1024, 12.5, 2.0, (True, True, True), False, -3.25, 0, 0.1, False),
    "dense_small_box": (2048, 11.95, 2.8, (True, True, True), True, 0.0, 0, 0.1, True),
}


def run_case(name, n, L, w, per, edges, lower, clusters, sigma, with_fields):
    """One step on the oracle; returns the fixture dict."""
    seed = sum(map(ord, name))
    d = synth(n, L, seed, lower=lower, clusters=clusters)
    rng = np.random.default_rng(seed + 1)
    disp = rng.normal(0.0, sigma, (3, n))
    box = po.Box(lower=(lower,) * 3, upper=(lower + L,) * 3, periodic=per, width=w,
                 edges_and_corners=edges)
    o = po.Oracle(box)
    cap = 27 * n + 1
    p = to_oracle(d, cap)
    fix = {"n": np.int64(n), "L": np.float64(L), "width": np.float64(w),
           "periodic": np.array(per, np.int32), "edges": np.int32(edges),
           "lower": np.float64(lower)}
    for k in F64 + I32:
        fix[f"in_{k}"] = d[k].copy()
    fix["in_disp"] = disp
    for a, k in enumerate(("x", "y", "z")):
        getattr(p, k)[:n] += disp[a]
    o.migrate(p, n)
    for k in ("x", "y", "z"):
        fix[f"migrated_{k}"] = getattr(p, k)[:n].copy()
    fields = []
    if with_fields:
        dens = np.zeros(cap)
        dens[:n] = rng.uniform(0, 2, n)
        vec = np.zeros((cap, 3))
        vec[:n] = rng.uniform(-1, 1, (n, 3))
        fix["in_density"] = dens[:n].copy()
        fix["in_vec"] = vec[:n].copy()
        fields = [po.Field(dens), po.Field(vec)]
    g = o.exchange_halo(p, n, fields=fields)
    fix["ghosts"] = np.int64(g)
    for k in F64 + I32:
        fix[f"halo_{k}"] = getattr(p, k)[:n + g].copy()
    fix["ghost_owners"] = o.ghost_owners()
    if with_fields:
        fix["halo_density"] = dens[:n + g].copy()
        fix["halo_vec"] = vec[:n + g].copy()
    forces = rng.uniform(-1, 1, (3, n + g))
    fix["in_forces"] = forces
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(p, k)[:n + g] = forces[a]
    o.reverse_halo(p, n)
    for k in ("fx", "fy", "fz"):
        fix[f"reversed_{k}"] = getattr(p, k)[:n].copy()
    if with_fields:
        contrib = rng.uniform(0, 1, n + g)
        fix["in_contrib"] = contrib
        dens[:n + g] = contrib
        o.reverse_exchange(n, [po.Field(dens, po.SUM)])
        fix["reversed_density"] = dens[:n].copy()
        vmax = rng.uniform(-1, 1, (n + g, 3))
        fix["in_vmax"] = vmax
        vec[:n + g] = vmax
        o.reverse_exchange(n, [po.Field(vec, po.MAX)])
        fix["reversed_vec_max"] = vec[:n].copy()
    return fix


def main():
    for name, spec in CASES.items():
        fix = run_case(name, *spec)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **fix)
        print(f"{path}: n={int(fix['n'])} ghosts={int(fix['ghosts'])}")


if __name__ == "__main__":
    main()
