"""Differentiation through the halo (65,536 particles, L = 40.3).  exchange_halo runs inside a
torch.autograd.Function whose backward is the reverse_halo kernel; a
Lennard-Jones energy over owned + ghost particles is differentiated, checked
by central finite differences (h = 1e-6) on 256 random particles, and
against forces from reverse_halo (GPU kernel bit-exact with the oracle's Host
reverse_halo on the same pair forces).  torch.autograd.gradcheck covers the
three Functions, and forward/reverse_exchange adjointness is checked on a
random FP64 3-vector field."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import assert_bits_equal, host, say, synth, to_device, to_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


def _setup(n, L, w, seed):
    from dsdecompx_amd import autograd as ag
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    d = synth(n, L, seed)
    p = to_device(d, 2 * n + 1024)
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    halo.migrate(p, n)
    g = ag.rebuild(halo, p, n)
    return d, p, halo, g


def test_gradcheck_functions(gpu):
    torch = gpu
    from dsdecompx_amd import autograd as ag
    say("gradcheck")
    n = 400
    d, p, halo, g = _setup(n, 7.0, 1.8, 31)
    assert g > 0
    gen = torch.Generator(device="cuda").manual_seed(3)
    v = torch.randn((n, 3), dtype=torch.float64, device="cuda", generator=gen, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: ag.halo_copy(t, halo), (v,), eps=1e-6, atol=1e-7,
                                    fast_mode=True)
    a = torch.randn((n + g, 3), dtype=torch.float64, device="cuda", generator=gen,
                    requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: ag.reverse_sum(t, halo), (a,), eps=1e-6, atol=1e-7,
                                    fast_mode=True)
    s = torch.randn((n, 5), dtype=torch.float64, device="cuda", generator=gen, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: ag.halo_copy(t, halo), (s,), eps=1e-6, atol=1e-7,
                                    fast_mode=True)
    pos = torch.stack([p.x[:n], p.y[:n], p.z[:n]], 1).clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: ag.halo_positions(t, halo, p), (pos,), eps=1e-6,
                                    atol=1e-7, fast_mode=True)
    # a rebuild invalidates graphs built on the previous plan
    out = ag.halo_copy(v, halo)
    ag.rebuild(halo, p, n)
    with pytest.raises(RuntimeError):
        out.sum().backward()


def test_lj_forces_65k(gpu):
    torch = gpu
    from dsdecompx_amd import autograd as ag
    from tests import lj
    n, L, w, rc = 65536, 40.3, 2.8, 2.5
    say("LJ", n)
    d, p, halo, g = _setup(n, L, w, 404)
    pos = torch.stack([p.x[:n], p.y[:n], p.z[:n]], 1).clone().requires_grad_(True)
    pos_all = ag.halo_positions(pos, halo, p)
    pairs = lj.pair_list(pos_all, n, rc)
    E = lj.energy(pos_all, pairs, rc)
    E.backward()
    grad = pos.grad.detach()

    # forces from reverse_halo: GPU kernel vs the oracle's Host reverse_halo
    f_all = lj.pair_forces(pos_all.detach(), pairs, rc)
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(p, k)[:n + g] = f_all[:, a]
    ref = to_oracle({k: host(getattr(p, k)[:n]) for k in
                     ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "charge",
                      "type", "global_id")}, n + g)
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
    assert ora.exchange_halo(ref, n) == g
    fa = host(f_all)
    ref.fx[:n + g], ref.fy[:n + g], ref.fz[:n + g] = fa[:, 0], fa[:, 1], fa[:, 2]
    ora.reverse_halo(ref, n)
    halo.reverse_halo(p, n)
    F = np.stack([host(p.fx[:n]), host(p.fy[:n]), host(p.fz[:n])], 1)
    for a, k in enumerate(("fx", "fy", "fz")):
        assert_bits_equal(F[:, a], getattr(ref, k)[:n], f"reverse_halo {k}")
    G = host(grad)
    scale = np.abs(F).max()
    assert np.abs(-G - F).max() <= 1e-10 * scale, "autograd force differs from reverse_halo force"

    # central finite differences on 256 random particles (local energy)
    owners = halo.ghost_owners(g).long()
    rng = np.random.default_rng(7)
    sample = rng.choice(n, 256, replace=False)
    i_all, j_all = pairs
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        base = pos.detach().clone()
        for k in sample:
            images = torch.nonzero(owners == int(k)).flatten() + n
            members = torch.cat([torch.tensor([int(k)], device="cuda"), images])
            mask = torch.isin(i_all, members) | torch.isin(j_all, members)
            local = (i_all[mask], j_all[mask])
            for a in range(3):
                q = base.clone()
                q[k, a] += h
                ep = lj.energy(ag.HaloPositions.apply(q, halo, p), local, rc)
                q[k, a] -= 2 * h
                em = lj.energy(ag.HaloPositions.apply(q, halo, p), local, rc)
                fd = float((ep - em) / (2 * h))
                err = abs(fd - float(G[k, a])) / (1.0 + abs(float(G[k, a])))
                worst = max(worst, err)
    assert worst < 1e-6, f"finite differences disagree: {worst}"


def test_forward_reverse_adjoint_vec3(gpu):
    torch = gpu
    from dsdecompx_amd import autograd as ag
    n, L, w = 65536, 40.3, 2.8
    say("adjointness", n)
    d, p, halo, g = _setup(n, L, w, 405)
    gen = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn((n, 3), dtype=torch.float64, device="cuda", generator=gen)
    y = torch.randn((g, 3), dtype=torch.float64, device="cuda", generator=gen)
    fx = ag.halo_copy(x, halo)
    s1 = float((fx[n:] * y).sum())
    yy = torch.zeros((n + g, 3), dtype=torch.float64, device="cuda")
    yy[n:] = y
    rv = ag.reverse_sum(yy, halo)
    s2 = float((x * rv).sum())
    assert abs(s1 - s2) <= 1e-10 * (1.0 + abs(s1))
    assert torch.equal(fx[:n], x)
