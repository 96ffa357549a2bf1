"""The batched form (khrr_accel_batched): many independent systems of uneven N
in one persistent launch.  Each system's result must equal the plain
single-system launch bit for bit, whatever else is in the batch: for
n >= 16384 the pair-symmetric launch (symmetric 1, i_per_lane 8), below it
the one-sided launch with j_split min(32, ceil(n/1024)), j_tile 256.  Results
must also match the FP64 oracle within the gravity tolerance (median <= 1e-6,
p99 <= 1e-5)."""
import numpy as np
import pytest

import ics
import oracle as orc
from helpers import err_stats, rel_err, to_dev

pytestmark = pytest.mark.gpu


def _population(sizes, seed0=100000):
    parts = []
    for k, n in enumerate(sizes):
        pos, vel = ics.plummer(int(n), 1.0, 0.3333, 1.0, seed=seed0 + k)
        eps = 0.004 * (max(n, 1) / 16384.0) ** (-1.0 / 3.0)
        parts.append((pos, np.full(n, 1.0 / max(n, 1)), np.full(n, 0.5 * eps * eps)))
    off = np.concatenate([[0], np.cumsum([p[0].shape[1] for p in parts])]).astype(np.int32)
    pos = np.concatenate([p[0] for p in parts], axis=1).astype(np.float32)
    m = np.concatenate([p[1] for p in parts]).astype(np.float32)
    h = np.concatenate([p[2] for p in parts]).astype(np.float32)
    return off, pos, m, h


def _run_batched(ctx, off, pos, m, h, tune=None, phi=False, compensated=False):
    import torch
    n = pos.shape[1]
    acc = torch.zeros((3, n), device="cuda:0")
    ph = torch.zeros(n, device="cuda:0") if phi else None
    ctx.accel_batched(to_dev(off, torch.int32), off.tolist(), to_dev(pos, torch.float32),
                      to_dev(h, torch.float32), to_dev(m, torch.float32), acc=acc, phi=ph,
                      tune=tune, compensated=compensated)
    torch.cuda.synchronize()
    return acc.cpu().numpy(), (ph.cpu().numpy() if phi else None)


def _plain_tune(n, j_tile=256):
    import khrr
    if n >= 16384:
        return khrr.Tune(8, 4, 1, 256, 1)
    return khrr.Tune(4, 4, max(1, min(32, -(-n // 1024))), j_tile, 0)


def _run_plain(ctx, pos, m, h, phi=False, compensated=False, j_tile=256):
    import torch
    n = pos.shape[1]
    acc = torch.zeros((3, n), device="cuda:0")
    ph = torch.zeros(n, device="cuda:0") if phi else None
    p = to_dev(pos, torch.float32)
    hh = to_dev(h, torch.float32)
    ctx.accel(p, hh, p, to_dev(m, torch.float32), hh, acc=acc, phi=ph,
              tune=_plain_tune(n, j_tile), compensated=compensated)
    torch.cuda.synchronize()
    return acc.cpu().numpy(), (ph.cpu().numpy() if phi else None)


def test_batched_equals_plain_per_system(ctx):
    sizes = [1024, 3072, 7168, 2048, 1024, 5120, 64, 1]
    off, pos, m, h = _population(sizes)
    acc, phi = _run_batched(ctx, off, pos, m, h, phi=True)
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        pa, pp = _run_plain(ctx, pos[:, a:b], m[a:b], h[a:b], phi=True)
        np.testing.assert_array_equal(acc[:, a:b], pa, err_msg=f"system {s}")
        np.testing.assert_array_equal(phi[a:b], pp, err_msg=f"system {s}")


@pytest.mark.parametrize("ipl", [2, 4, 8])
def test_batched_oracle_and_shapes(ctx, ipl):
    import khrr
    sizes = [4096, 1024, 2048, 3072]
    off, pos, m, h = _population(sizes, seed0=200)
    acc, _ = _run_batched(ctx, off, pos, m, h, tune=khrr.Tune(ipl, 2, 1, 256))
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        ref, _ = orc.accel(pos[:, a:b], h[a:b], pos[:, a:b], m[a:b], h[a:b], want_phi=False)
        st = err_stats(rel_err(acc[:, a:b], ref))
        assert st["median"] <= 1e-6 and st["p99"] <= 1e-5, (s, st)


def test_nonuniform_galaxy_equals_plain(ctx):
    # stars and halo bodies of different mass in one system: both forms take
    # the same per-tile paths
    from helpers import galaxy_system
    pos, _, mass, h = galaxy_system(1000, 3000, seed=9, eps_s=0.004, eps_d=0.006)
    p32, m32, h32 = pos.astype(np.float32), mass.astype(np.float32), h.astype(np.float32)
    off = np.array([0, 1500, p32.shape[1]], np.int32)
    acc, phi = _run_batched(ctx, off, p32, m32, h32, phi=True)
    for s in range(2):
        a, b = off[s], off[s + 1]
        pa, pp = _run_plain(ctx, p32[:, a:b].copy(), m32[a:b].copy(), h32[a:b].copy(), phi=True)
        np.testing.assert_array_equal(acc[:, a:b], pa)
        np.testing.assert_array_equal(phi[a:b], pp)


def test_alone_equals_in_the_mix(ctx):
    # the same system alone and embedded among others (different position
    # in the work list, different neighbours): identical bits
    off, pos, m, h = _population([2048, 5120, 1024], seed0=300)
    acc_mix, _ = _run_batched(ctx, off, pos, m, h)
    a, b = off[1], off[2]
    solo_off = np.array([0, b - a], np.int32)
    acc_solo, _ = _run_batched(ctx, solo_off, pos[:, a:b].copy(), m[a:b].copy(), h[a:b].copy())
    np.testing.assert_array_equal(acc_mix[:, a:b], acc_solo)


def test_composition_change_and_empty_systems(ctx):
    off, pos, m, h = _population([1024, 0, 2048], seed0=400)
    acc1, _ = _run_batched(ctx, off, pos, m, h)
    off2, pos2, m2, h2 = _population([3072, 1024], seed0=500)
    acc2, _ = _run_batched(ctx, off2, pos2, m2, h2)
    acc1b, _ = _run_batched(ctx, off, pos, m, h)
    np.testing.assert_array_equal(acc1, acc1b)
    assert np.isfinite(acc2).all()


@pytest.mark.parametrize("compensated", [False, True])
def test_large_systems_equal_plain(ctx, compensated):
    # n < 16384: j-range cut into ~1024-body pieces summed in piece order;
    # n >= 16384: pair-symmetric superblock items; both agree exactly with
    # the plain launch of the same shape
    sizes = [40000, 1000, 16385, 70000, 16384, 32768, 32767]
    off, pos, m, h = _population(sizes, seed0=700)
    acc, phi = _run_batched(ctx, off, pos, m, h, phi=True, compensated=compensated)
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        pa, pp = _run_plain(ctx, pos[:, a:b].copy(), m[a:b].copy(), h[a:b].copy(), phi=True,
                            compensated=compensated)
        np.testing.assert_array_equal(acc[:, a:b], pa, err_msg=f"system {s}")
        np.testing.assert_array_equal(phi[a:b], pp, err_msg=f"system {s}")
    a, b = off[0], off[1]
    rows = np.arange(0, b - a, 97)
    ref, _ = orc.accel_rows(rows, pos[:, a:b], h[a:b], m[a:b])
    st = err_stats(rel_err(acc[:, a + rows], ref))
    assert st["median"] <= 1e-6 and st["p99"] <= 1e-5, st


def test_bad_offsets_rejected(ctx):
    import khrr
    import torch
    pos = torch.zeros((3, 4), device="cuda:0")
    h = torch.zeros(4, device="cuda:0")
    with pytest.raises(khrr.KhrrError, match="non-decreasing"):
        ctx.accel_batched(torch.tensor([0, 3, 2, 4], dtype=torch.int32, device="cuda:0"),
                          [0, 3, 2, 4], pos, h, h, acc=torch.zeros((3, 4), device="cuda:0"))


def test_autotuners_return_valid_shapes():
    # a context of its own: the tuned shapes it stores (a tuned j-tile
    # changes the one-sided sums) stay out of the other tests' context
    import khrr
    ctx = khrr.Context(0)
    try:
        t, ms = ctx.tune_gravity(20000)
        assert t.i_per_lane in (1, 2, 4, 8) and 1 <= t.waves_per_block <= 4 and ms > 0
        off, pos, m, h = _population([1024, 2048, 4096], seed0=600)
        tb, msb = ctx.tune_batched(off.tolist())
        assert tb.i_per_lane in (2, 4, 8) and tb.j_split == 1 and msb > 0
        assert tb.j_tile in (128, 256, 512, 1024)
        # the tuned shape is what a NULL tune now uses; each system equals
        # its plain launch with the same j-tile bit for bit
        acc, _ = _run_batched(ctx, off, pos, m, h)
        for s in range(3):
            a, b = off[s], off[s + 1]
            pa, _ = _run_plain(ctx, pos[:, a:b], m[a:b], h[a:b], j_tile=tb.j_tile)
            np.testing.assert_array_equal(acc[:, a:b], pa)
    finally:
        ctx.close()


def test_graph_replays_equal_eager(ctx):
    # KHRR_GRAPH: the first call runs eagerly and captures; later calls with
    # the same composition and pointers replay the graph.  Every replay,
    # with eager work and synchronisations in between and with the inputs
    # changed in place, must equal the eager result bit for bit.
    import torch
    sizes = [40000, 3072, 17000, 1024, 70000]
    off, pos, m, h = _population(sizes, seed0=800)
    n = pos.shape[1]
    offd = to_dev(off, torch.int32)
    P, H, M = to_dev(pos, torch.float32), to_dev(h, torch.float32), to_dev(m, torch.float32)
    acc_g = torch.zeros((3, n), device="cuda:0")
    phi_g = torch.zeros(n, device="cuda:0")
    rng = np.random.default_rng(1)
    for rep in range(5):
        if rep >= 2:  # new positions in the same buffer
            P.add_(to_dev(rng.normal(scale=1e-3, size=(3, n)), torch.float32))
        acc_g.fill_(7.0)
        ctx.accel_batched(offd, off.tolist(), P, H, M, acc=acc_g, phi=phi_g, graph=True)
        torch.cuda.synchronize()
        got_a, got_p = acc_g.cpu().numpy(), phi_g.cpu().numpy()
        # eager work in between, then the eager reference on the same inputs
        acc_e = torch.zeros((3, n), device="cuda:0")
        phi_e = torch.zeros(n, device="cuda:0")
        ctx.accel_batched(offd, off.tolist(), P, H, M, acc=acc_e, phi=phi_e)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got_a, acc_e.cpu().numpy(), err_msg=f"replay {rep}")
        np.testing.assert_array_equal(got_p, phi_e.cpu().numpy(), err_msg=f"replay {rep}")


def test_graph_with_compositions_changing(ctx):
    # more compositions than cached graphs: eviction and recapture stay exact
    import torch
    for k in range(6):
        sizes = [1024 * (k + 1), 2048, 33000 + 1000 * (k % 2)]
        off, pos, m, h = _population(sizes, seed0=900 + k)
        n = pos.shape[1]
        offd = to_dev(off, torch.int32)
        P, H, M = to_dev(pos, torch.float32), to_dev(h, torch.float32), to_dev(m, torch.float32)
        outs = []
        for graph in (True, True, False):
            acc = torch.zeros((3, n), device="cuda:0")
            ctx.accel_batched(offd, off.tolist(), P, H, M, acc=acc, graph=graph)
            torch.cuda.synchronize()
            outs.append(acc.cpu().numpy())
        np.testing.assert_array_equal(outs[0], outs[2])
        np.testing.assert_array_equal(outs[1], outs[2])
