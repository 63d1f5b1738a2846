"""GPU parity of the HIP path (through the C ABI) against the CPU oracle on
identical seeded inputs: ghost lists, positions and every particle array
bit-exact; reverse sums bit-exact (same per-owner order); migrate wrap
bit-exact."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.gpu_util import (F64, I32, assert_bits_equal, compare_particles, host, say, synth,
                            to_device, to_oracle)

from tests.gpu_util import box_ghost_writes  # noqa: F401  (autouse: every ghost-write kernel)
pytestmark = pytest.mark.gpu

CASES = [
    # name, n, L, width, periodic, edges, lower
    ("empty", 0, 8.0, 1.0, (True, True, True), True, 0.0),
    ("one", 1, 8.0, 1.0, (True, True, True), True, 0.0),
    ("small", 1000, 10.0, 1.5, (True, True, True), True, 0.0),
    ("ragged_chunk", 4099, 16.0, 2.8, (True, True, True), True, 0.0),
    ("uniform_32k", 32768, 32.0, 2.8, (True, True, True), True, 0.0),
    ("z_nonperiodic", 20000, 27.0, 2.8, (True, True, False), True, 0.0),
    ("faces_only", 20000, 27.0, 2.8, (True, True, True), False, 0.0),
    ("shifted_box", 15000, 24.6, 2.8, (True, True, True), True, -7.3),
    ("wide_halo", 3000, 5.0, 3.1, (True, True, True), True, 0.0),
    ("dense_small", 2048, 11.95, 2.8, (True, True, True), True, 0.0),
]


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dsdecompx_amd import _abi
    _abi.lib()
    # the device path requires an explicit (non-null) stream, like the reference
    torch.cuda.set_stream(torch.cuda.Stream())
    return torch


def _run_build(case, seed=1, fields=False):
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    name, n, L, w, per, edges, lower = case
    upper = tuple(lower + L for _ in range(3))
    box_o = po.Box(lower=(lower,) * 3, upper=upper, periodic=per, width=w, edges_and_corners=edges)
    box_d = BoxConfig(lower=(lower,) * 3, upper=upper, periodic=per, width=w,
                      edges_and_corners=edges)
    d = synth(n, L, seed, lower=lower)
    cap = max(64, n * 27 + 1) if n < 5000 else n * 3
    ora = po.Oracle(box_o)
    op = to_oracle(d, cap)
    g_ref = ora.exchange_halo(op, n)
    dev = to_device(d, max(1, n + 16))   # deliberately small: forces the regrow path
    halo = LoopbackHalo(box_d)
    g = halo.exchange_halo(dev, n)
    return ora, op, g_ref, halo, dev, g, n


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_halo_build_bit_exact(gpu, case):
    say("halo build", case)
    ora, op, g_ref, halo, dev, g, n = _run_build(case)
    assert g == g_ref
    compare_particles(dev, op, n + g, case[0])
    send, recv, g2, owners, ov = halo.query()
    nb = ora.nb
    ref_send = dict(zip(nb.slot.tolist(), ora.plan.send_counts(nb).tolist()))
    assert [send[s] for s in range(26)] == [ref_send.get(s, 0) for s in range(26)]
    assert recv == [send[25 - m] for m in range(26)]
    assert g2 == g and ov == 0
    own = host(halo.ghost_owners(g))
    assert_bits_equal(own, ora.ghost_owners(), "ghost owners")
    assert owners == len(np.unique(own))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_reverse_halo_bit_exact(gpu, case):
    import torch
    say("reverse halo", case)
    ora, op, g_ref, halo, dev, g, n = _run_build(case, seed=2)
    rng = np.random.default_rng(5)
    for k in ("fx", "fy", "fz"):
        vals = rng.uniform(-1, 1, n + g)
        getattr(op, k)[:n + g] = vals
        getattr(dev, k)[:n + g] = torch.from_numpy(vals).to(dev.device)
    ora.reverse_halo(op, n)
    halo.reverse_halo(dev, n)
    for k in ("fx", "fy", "fz"):
        assert_bits_equal(host(getattr(dev, k)[:n + g]), getattr(op, k)[:n + g], k)


def test_upper_cut_tie_known_answer(gpu):
    """Box [9.8, 53.6), w = 1.1: the local domain's upper cut is
    lower + extent = 53.599999999999994 (src/topology.cpp:66-72), so the
    high threshold is 52.49999999999999 and particles at exactly 52.5 are in
    the high slab.  Given cuts (local=) use the domain's own bounds instead,
    where 52.5 is on the threshold.  Bit-exact against the oracle both ways,
    with the known ghost counts of the tie particles."""
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    say("upper cut tie KAT")
    lo, up, w = 9.8, 53.6, 1.1
    n = 3000
    d = synth(n, up - lo, 21, lower=lo)
    mid = 30.0
    ties = [(52.5, mid, mid), (mid, 52.5, mid), (mid, mid, 52.5), (52.5, 52.5, mid),
            (52.49999999999999, mid, mid), (10.9, mid, mid)]
    for k, (x, y, z) in enumerate(ties):
        d["x"][k], d["y"][k], d["z"][k] = x, y, z
    for local, tie_images in ((None, [1, 1, 1, 3, 0, 0]), (((lo,) * 3, (up,) * 3), [0, 0, 0, 0, 0, 0])):
        box_o = po.Box(lower=(lo,) * 3, upper=(up,) * 3, width=w, local=local)
        box_d = BoxConfig(lower=(lo,) * 3, upper=(up,) * 3, width=w, local=local)
        ora = po.Oracle(box_o)
        op = to_oracle(d, 3 * n)
        g_ref = ora.exchange_halo(op, n)
        dev = to_device(d, 3 * n)
        halo = LoopbackHalo(box_d)
        assert halo.exchange_halo(dev, n) == g_ref
        compare_particles(dev, op, n + g_ref, f"tie box local={local}")
        own = host(halo.ghost_owners(g_ref))
        assert_bits_equal(own, ora.ghost_owners(), "ghost owners")
        assert [int((own == k).sum()) for k in range(len(ties))] == tie_images


def test_reverse_order_known_answer(gpu):
    """Catastrophic cancellation makes only ascending send-slot order give 4.75."""
    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    say("reverse order KAT")
    d = synth(1, 10.0, 0)
    d["x"][:], d["y"][:], d["z"][:] = 0.2, 0.3, 0.4
    dev = to_device(d, 16)
    halo = LoopbackHalo(BoxConfig(upper=(10.0,) * 3, width=1.0))
    assert halo.exchange_halo(dev, 1) == 7
    blocks = [13, 15, 16, 21, 22, 24, 25]
    value = {0: 1e16, 1: 1.0, 3: -1e16, 4: 1.0, 9: 0.5, 10: 0.25, 12: 3.0}
    fx = np.zeros(8)
    fx[0] = 1.0
    for s, v in value.items():
        fx[1 + blocks.index(25 - s)] = v
    dev.fx[:8] = torch.from_numpy(fx).cuda()
    halo.reverse_halo(dev, 1)
    assert host(dev.fx[:1])[0] == 4.75


def test_migrate_wrap_bit_exact(gpu):
    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    say("migrate wrap")
    rng = np.random.default_rng(17)
    for lower, L, per in [(0.0, 32.0, (True, True, True)), (-3.7, 15.8, (True, False, True)),
                          (0.0, 256.0, (True, True, False))]:
        n = 200000
        x = rng.uniform(lower - 3 * L, lower + 4 * L, n)
        x[:8] = [-1e-17, lower + L, lower, -0.0, lower + L + 1e-12, np.nextafter(lower, -1),
                 np.inf, np.nan]
        y = np.roll(x, 3)
        z = np.roll(x, 7)
        box_o = po.Box(lower=(lower,) * 3, upper=(lower + L,) * 3, periodic=per)
        ox, oy, oz = x.copy(), y.copy(), z.copy()
        po.lib().dsdxo_migrate_wrap(__import__("ctypes").byref(box_o.c()), n, po._dp(ox),
                                    po._dp(oy), po._dp(oz))
        d = synth(n, L, 1)
        d["x"], d["y"], d["z"] = x, y, z
        dev = to_device(d, n)
        halo = LoopbackHalo(BoxConfig(lower=(lower,) * 3, upper=(lower + L,) * 3, periodic=per))
        halo.migrate(dev, n)
        for k, ref in (("x", ox), ("y", oy), ("z", oz)):
            got = host(getattr(dev, k)[:n])
            nan = np.isnan(ref)
            assert np.array_equal(np.isnan(got), nan)
            assert_bits_equal(got[~nan], ref[~nan], f"wrap {k} lower={lower}")


FIELD_TYPES = [np.int32, np.int64, np.uint32, np.uint64, np.float32, np.float64]


def test_fields_forward_and_reverse(gpu):
    import torch
    from dsdecompx_amd import _abi
    say("fields")
    case = ("fields", 6000, 14.0, 2.8, (True, True, True), True, 0.0)
    ora, op, g_ref, halo, dev, g, n = _run_build(case, seed=3)
    rng = np.random.default_rng(8)
    for dt in FIELD_TYPES:
        for comps in (1, 3):
            for op_code in (_abi.SUM, _abi.MIN, _abi.MAX, _abi.REPLACE):
                shape = (n + g,) if comps == 1 else (n + g, comps)
                if np.issubdtype(dt, np.integer):
                    info = np.iinfo(dt)
                    vals = rng.integers(max(info.min, -2**40), min(info.max, 2**40), shape,
                                        dtype=np.int64 if dt != np.uint64 else np.uint64).astype(dt)
                else:
                    vals = rng.standard_normal(shape).astype(dt)
                ref = vals.copy()
                ora.reverse_exchange(n, [po.Field(ref, op_code)])
                t = torch.from_numpy(vals.copy()).cuda()
                halo.reverse_exchange([(t, op_code)], n)
                assert_bits_equal(host(t).reshape(-1), ref.reshape(-1),
                                  f"reverse {np.dtype(dt).name}x{comps} op{op_code}")
            # forward: ghosts get exact owner copies, owners untouched
            shape = (n + g,) if comps == 1 else (n + g, comps)
            vals = (rng.integers(0, 1000, shape)).astype(dt)
            ref = vals.copy()
            ora.forward_exchange(n, [po.Field(ref)])
            t = torch.from_numpy(vals.copy()).cuda()
            halo.forward_exchange([t], n)
            assert_bits_equal(host(t).reshape(-1), ref.reshape(-1), f"forward {np.dtype(dt).name}")


def test_fields_travel_with_halo_build(gpu):
    """Registered fields are copied into the ghosts by the rebuild itself
    (src/halo.cpp:120-147, :874-890)."""
    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    say("fields in halo build")
    n, L = 5000, 13.0
    d = synth(n, L, 4)
    rng = np.random.default_rng(4)
    dens = np.zeros(n * 3)
    dens[:n] = rng.uniform(0, 2, n)
    vec = np.zeros((n * 3, 3))
    vec[:n] = rng.uniform(-1, 1, (n, 3))
    box_o = po.Box(upper=(L,) * 3, width=2.8)
    ora = po.Oracle(box_o)
    op = to_oracle(d, n * 3)
    rd, rv = dens.copy(), vec.copy()
    g = ora.exchange_halo(op, n, fields=[po.Field(rd), po.Field(rv)])
    dev = to_device(d, n * 3)
    td, tv = torch.from_numpy(dens.copy()).cuda(), torch.from_numpy(vec.copy()).cuda()
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=2.8))
    assert halo.exchange_halo(dev, n, fields=[td, tv]) == g
    assert_bits_equal(host(td)[:n + g], rd[:n + g], "density")
    assert_bits_equal(host(tv)[:n + g].reshape(-1), rv[:n + g].reshape(-1), "vec3")


def test_refresh_matches_rewrite(gpu):
    """A reused plan re-reads every owner (positions_only or the whole
    record); the oracle rewrites the ghosts of its stored plan the same way."""
    import torch
    say("refresh")
    case = ("refresh", 8000, 15.0, 2.8, (True, True, True), True, 0.0)
    ora, op, g_ref, halo, dev, g, n = _run_build(case, seed=6)
    rng = np.random.default_rng(6)
    for full in (False, True):
        dx = rng.normal(0, 0.05, (3, n))
        for a, k in enumerate(("x", "y", "z")):
            getattr(op, k)[:n] += dx[a]
            getattr(dev, k)[:n] = torch.from_numpy(getattr(op, k)[:n].copy()).cuda()
        if full:
            op.vx[:n] += 1.0
            dev.vx[:n] = torch.from_numpy(op.vx[:n].copy()).cuda()
        before = {k: host(getattr(dev, k)[:n + g]).copy() for k in F64 + I32}
        halo.refresh(dev, n, positions_only=not full)
        ref = to_oracle({k: getattr(op, k)[:n].copy() for k in F64 + I32}, n + g)
        for k in F64 + I32:
            getattr(ref, k)[n:n + g] = before[k][n:]
        ora.write_ghosts(ref, n)
        if not full:   # positions_only: only x, y, z and zero forces change
            for k in ("vx", "vy", "vz", "mass", "charge", "type", "global_id"):
                getattr(ref, k)[n:n + g] = before[k][n:]
        compare_particles(dev, ref, n + g, f"refresh full={full}")


def test_multi_step_uniform_32k(gpu):
    """32,768 particles in L = 32, 10 steps: displace N(0,0.05), migrate, exchange_halo, forces U[-1,1]
    on owned+ghosts, reverse_halo; every array bit-exact after every step."""
    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    n, L, w = 32768, 32.0, 2.8
    d = synth(n, L, 100)
    box_o = po.Box(upper=(L,) * 3, width=w)
    ora = po.Oracle(box_o)
    op = to_oracle(d, n * 3)
    dev = to_device(d, n * 3)
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    rng = np.random.default_rng(1234)
    for step in range(10):
        say("32k step", step)
        disp = rng.normal(0, 0.05, (3, n))
        for a, k in enumerate(("x", "y", "z")):
            getattr(op, k)[:n] += disp[a]
            getattr(dev, k)[:n] += torch.from_numpy(disp[a]).cuda()
        ora.migrate(op, n)
        halo.migrate(dev, n)
        g = ora.exchange_halo(op, n)
        assert halo.exchange_halo(dev, n) == g
        f = rng.uniform(-1, 1, (3, n + g))
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(op, k)[:n + g] = f[a]
            getattr(dev, k)[:n + g] = torch.from_numpy(f[a]).cuda()
        ora.reverse_halo(op, n)
        halo.reverse_halo(dev, n)
        compare_particles(dev, op, n + g, f"step {step}")


def test_one_plan_across_sizes(gpu):
    """One plan rebuilt for 1,500, then 60,000, then 1,500 particles in the
    same box.  The owner-major passes size their launch from the previous
    build's owner count, so the second build's copy runs with far fewer waves
    than owner tiles (each wave walks several) and the third with more (the
    extra waves have no tile); every array and reverse sum bit-exact."""
    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    L, w = 32.0, 2.8
    box_o = po.Box(upper=(L,) * 3, width=w)
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    rng = np.random.default_rng(321)
    for step, n in enumerate((1500, 60000, 1500)):
        say("sizes", n)
        d = synth(n, L, 400 + step)
        ora = po.Oracle(box_o)
        op = to_oracle(d, n * 3)
        dev = to_device(d, n * 3)
        g = ora.exchange_halo(op, n)
        assert halo.exchange_halo(dev, n) == g
        f = rng.uniform(-1, 1, (3, n + g))
        for a, k in enumerate(("fx", "fy", "fz")):
            getattr(op, k)[:n + g] = f[a]
            getattr(dev, k)[:n + g] = torch.from_numpy(f[a]).cuda()
        ora.reverse_halo(op, n)
        halo.reverse_halo(dev, n)
        compare_particles(dev, op, n + g, f"sizes n={n}")


@pytest.mark.parametrize("items,waves,descending", [(16, 4, 127), (4, 2, 0), (8, 8, 24)])
def test_launch_shapes_and_walk_orders(gpu, monkeypatch, items, waves, descending):
    """Launch shape and walk order (DSDX_ITEMS, DSDX_WAVES, DSDX_DESCENDING,
    read when a plan is made) change no result: migrate, build, reverse and
    a full refresh bit-exact with the oracle with every pass walking last to
    first, and at other chunk sizes."""
    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    monkeypatch.setenv("DSDX_ITEMS", str(items))
    monkeypatch.setenv("DSDX_WAVES", str(waves))
    monkeypatch.setenv("DSDX_DESCENDING", str(descending))
    n, L, w = 20000, 27.0, 2.8
    say("shape", items, waves, descending)
    d = synth(n, L, 500 + items)
    rng = np.random.default_rng(501)
    for k in ("x", "y", "z"):
        d[k] = d[k] + rng.normal(0, 0.3, n)
    ora = po.Oracle(po.Box(upper=(L,) * 3, width=w))
    op = to_oracle(d, n * 3)
    dev = to_device(d, n * 3)
    halo = LoopbackHalo(BoxConfig(upper=(L,) * 3, width=w))
    ora.migrate(op, n)
    halo.migrate(dev, n)
    g = ora.exchange_halo(op, n)
    assert halo.exchange_halo(dev, n) == g
    f = rng.uniform(-1, 1, (3, n + g))
    for a, k in enumerate(("fx", "fy", "fz")):
        getattr(op, k)[:n + g] = f[a]
        getattr(dev, k)[:n + g] = torch.from_numpy(f[a]).cuda()
    ora.reverse_halo(op, n)
    halo.reverse_halo(dev, n)
    compare_particles(dev, op, n + g, f"shape {items}x{waves} order {descending}")
    # a full refresh rewrites every ghost from its owner, as a rebuild does
    op.vx[:n] += 1.0
    dev.vx[:n] = torch.from_numpy(op.vx[:n].copy()).cuda()
    halo.refresh(dev, n, positions_only=False)
    ora.write_ghosts(op, n)
    compare_particles(dev, op, n + g, f"refresh {items}x{waves} order {descending}")


@pytest.mark.parametrize("fused_order", [0, 64], ids=["first_to_last", "last_to_first"])
@pytest.mark.parametrize("case", [c for c in CASES if c[0] in ("uniform_32k", "z_nonperiodic",
                                                                 "shifted_box", "empty")],
                         ids=lambda c: c[0])
def test_fused_migrate_build_equals_separate(gpu, monkeypatch, case, fused_order):
    """dsdx_migrate_halo_build == dsdx_migrate_wrap + dsdx_halo_build, bit for bit,
    with the fused pass walking its chunks either way (DSDX_DESCENDING bit 64)."""
    import torch
    from dsdecompx_amd.halo import BoxConfig, LoopbackHalo
    monkeypatch.setenv("DSDX_DESCENDING", str(1 | 16 | fused_order))
    name, n, L, w, per, edges, lower = case
    say("fused", name, fused_order)
    upper = tuple(lower + L for _ in range(3))
    box = BoxConfig(lower=(lower,) * 3, upper=upper, periodic=per, width=w, edges_and_corners=edges)
    d = synth(n, L, 77, lower=lower)
    rng = np.random.default_rng(78)
    for k in ("x", "y", "z"):
        d[k] = d[k] + rng.normal(0, 0.7, n)
    d["x"][:min(n, 3)] = [-1e-17, lower + L, lower - L][:min(n, 3)]
    a = to_device(d, max(64, 4 * n))
    b = to_device(d, max(64, 4 * n))
    h1, h2 = LoopbackHalo(box), LoopbackHalo(box)
    h1.migrate(a, n)
    g1 = h1.exchange_halo(a, n)
    g2 = h2.migrate_exchange_halo(b, n)
    assert g1 == g2
    for k in F64 + I32:
        assert_bits_equal(host(getattr(a, k)[:n + g1]), host(getattr(b, k)[:n + g1]), f"fused {k}")
