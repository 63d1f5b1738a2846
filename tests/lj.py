"""Lennard-Jones pair energy and forces over owned + ghost particles (torch,
FP64): the application side of the differentiation test (65,536 particles).

E = 1/2 sum_{i owned} sum_{j owned or ghost, j != i, r_ij < rc} phi(r_ij),
phi(r) = 4 eps ((sigma/r)^12 - (sigma/r)^6) - phi(rc)  (energy-shifted so the
energy is continuous at rc; forces are those of plain LJ, cf. the reference's
examples/applications/md/common.h:323-336).

Pairs are found once with a cell list and kept fixed, so E is a smooth
function of the positions for finite differences.
"""
from __future__ import annotations

import torch


def phi(r2, eps=1.0, sigma=1.0, rc=2.5):
    s6 = (sigma * sigma / r2) ** 3
    rc6 = (sigma / rc) ** 6
    return 4.0 * eps * (s6 * s6 - s6) - 4.0 * eps * (rc6 * rc6 - rc6)


def pair_list(pos_all: torch.Tensor, owned: int, rc: float, skin: float = 0.0):
    """Ordered pairs (i owned, j any, j != i) with r_ij < rc + skin."""
    device = pos_all.device
    p = pos_all.detach()
    cut = rc + skin
    lo = p.min(0).values
    ncell = torch.clamp(((p.max(0).values - lo) / cut).floor().long() + 1, min=1)
    cell3 = torch.clamp(((p - lo) / cut).floor().long(), min=torch.zeros(3, dtype=torch.long, device=device),
                        max=ncell - 1)
    cid = (cell3[:, 0] * ncell[1] + cell3[:, 1]) * ncell[2] + cell3[:, 2]
    order = torch.argsort(cid)
    counts = torch.bincount(cid, minlength=int(ncell.prod()))
    start = torch.cumsum(counts, 0) - counts
    maxc = int(counts.max())
    ntot = int(ncell.prod())
    table = torch.full((ntot, maxc), -1, dtype=torch.long, device=device)
    rank = torch.arange(p.shape[0], device=device) - start[cid[order]]
    table[cid[order], rank] = order
    ii, jj = [], []
    own = torch.arange(owned, device=device)
    c_own = cell3[:owned]
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                nc = c_own + torch.tensor([dx, dy, dz], device=device)
                ok = ((nc >= 0) & (nc < ncell)).all(1)
                ncid = (nc[:, 0] * ncell[1] + nc[:, 1]) * ncell[2] + nc[:, 2]
                ncid = torch.where(ok, ncid, torch.zeros_like(ncid))
                cand = table[ncid]                       # [owned, maxc]
                cand = torch.where(ok[:, None], cand, torch.full_like(cand, -1))
                i = own[:, None].expand_as(cand)
                m = (cand >= 0) & (cand != i)
                i, j = i[m], cand[m]
                d = p[i] - p[j]
                keep = (d * d).sum(1) < cut * cut
                ii.append(i[keep])
                jj.append(j[keep])
    return torch.cat(ii), torch.cat(jj)


def energy(pos_all: torch.Tensor, pairs, rc=2.5):
    i, j = pairs
    d = pos_all[i] - pos_all[j]
    r2 = (d * d).sum(1)
    inside = r2 < rc * rc
    return 0.5 * torch.where(inside, phi(torch.where(inside, r2, torch.ones_like(r2)), rc=rc),
                             torch.zeros_like(r2)).sum()


def pair_forces(pos_all: torch.Tensor, pairs, rc=2.5):
    """Force on every particle (owned and ghost) from E: each ordered pair
    (i, j) contributes 1/2 of its pair force to i and the opposite to j; the
    ghosts' shares are returned to their owners by reverse_halo."""
    i, j = pairs
    d = pos_all[i] - pos_all[j]
    r2 = (d * d).sum(1)
    inside = r2 < rc * rc
    r2s = torch.where(inside, r2, torch.ones_like(r2))
    s6 = (1.0 / r2s) ** 3
    # -dphi/dr * (d / r) = 24 (2 s12 - s6) / r2 * d
    coef = torch.where(inside, 24.0 * (2.0 * s6 * s6 - s6) / r2s, torch.zeros_like(r2))
    f = 0.5 * coef[:, None] * d
    out = torch.zeros_like(pos_all)
    out.index_add_(0, i, f)
    out.index_add_(0, j, -f)
    return out
