"""The pair-symmetric gravity form (each unordered pair evaluated once,
khrr-nbody_amd/csrc/khrr_sym.hip) against the FP64 oracle (ACCEL/PHI,
nbody-kernel.cuh:43-142) and against the one-sided kernel.

Tolerances as the one-sided form: acceleration median <= 1e-6, p99 <= 1e-5
per particle; potential max relative <= 2e-6."""
import numpy as np
import pytest

import oracle as orc
from helpers import err_stats, galaxy_system, rel_err, to_dev

pytestmark = pytest.mark.gpu


def _run(ctx, p32, m32, h32, tune, phi=True, **kw):
    import torch
    pos = to_dev(p32, torch.float32)
    m = to_dev(m32, torch.float32)
    h = to_dev(h32, torch.float32)
    n = p32.shape[1]
    acc = torch.zeros((3, n), device="cuda:0")
    ph = torch.zeros(n, device="cuda:0") if phi else None
    ctx.accel(pos, h, pos, m, h, acc=acc, phi=ph, tune=tune, **kw)
    torch.cuda.synchronize()
    return acc.cpu().numpy().astype(np.float64), (ph.cpu().numpy().astype(np.float64)
                                                  if phi else None)


def _sym(ipl=4):
    import khrr
    return khrr.Tune(ipl, 4, 1, 256, 1)


@pytest.mark.parametrize("ipl", [4, 8])
def test_galaxy_uniform_path(ctx, ipl):
    # 3 stars superblocks + a ragged tail: uniform softening and mass per
    # superblock except the star/halo boundary block
    pos, _, mass, h = galaxy_system(6000, 9000, seed=31)
    p32, m32, h32 = pos.astype(np.float32), mass.astype(np.float32), h.astype(np.float32)
    acc, phi = _run(ctx, p32, m32, h32, _sym(ipl))
    ref_acc, ref_phi = orc.accel(p32, h32, p32, m32, h32)
    s = err_stats(rel_err(acc, ref_acc))
    print("symmetric", ipl, s)
    assert s["median"] <= 1e-6 and s["p99"] <= 1e-5
    assert np.max(np.abs(phi - ref_phi) / np.abs(ref_phi)) <= 2e-6


def test_generic_path_mixed_softening_and_mass(ctx):
    rng = np.random.default_rng(8)
    n = 9000
    p32 = rng.normal(size=(3, n)).astype(np.float32)
    m32 = rng.uniform(0.5, 2.0, n).astype(np.float32) / n
    h32 = rng.uniform(5e-6, 5e-5, n).astype(np.float32)
    acc, phi = _run(ctx, p32, m32, h32, _sym())
    ref_acc, ref_phi = orc.accel(p32, h32, p32, m32, h32)
    s = err_stats(rel_err(acc, ref_acc))
    assert s["median"] <= 1e-6 and s["p99"] <= 1e-5, s
    assert np.max(np.abs(phi - ref_phi) / np.abs(ref_phi)) <= 2e-6


def test_compensated_and_accumulate(ctx):
    import torch
    pos, _, mass, h = galaxy_system(4096, 8192, seed=2)
    p32, m32, h32 = pos.astype(np.float32), mass.astype(np.float32), h.astype(np.float32)
    acc_c, phi_c = _run(ctx, p32, m32, h32, _sym(), compensated=True)
    ref_acc, ref_phi = orc.accel(p32, h32, p32, m32, h32)
    s = err_stats(rel_err(acc_c, ref_acc))
    assert s["median"] <= 1e-6 and s["p99"] <= 1e-5, s
    # accumulate: acc += sum, phi -= sum
    n = p32.shape[1]
    a0 = np.random.default_rng(1).normal(size=(3, n)).astype(np.float32)
    acc = to_dev(a0, torch.float32)
    phi0 = torch.ones(n, device="cuda:0")
    P, M, H = (to_dev(v, torch.float32) for v in (p32, m32, h32))
    ctx.accel(P, H, P, M, H, acc=acc, phi=phi0, accumulate=True, tune=_sym())
    a1, p1 = _run(ctx, p32, m32, h32, _sym())
    torch.cuda.synchronize()
    np.testing.assert_allclose(acc.cpu().numpy(), a0 + a1, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(phi0.cpu().numpy(), 1.0 + p1, rtol=1e-6)


def test_deterministic_and_close_to_one_sided(ctx):
    import khrr
    pos, _, mass, h = galaxy_system(8192, 8192, seed=4)
    p32, m32, h32 = pos.astype(np.float32), mass.astype(np.float32), h.astype(np.float32)
    a1, f1 = _run(ctx, p32, m32, h32, _sym())
    a2, f2 = _run(ctx, p32, m32, h32, _sym())
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(f1, f2)
    a3, f3 = _run(ctx, p32, m32, h32, khrr.Tune(2, 4, 4, 256, 0))
    s = err_stats(rel_err(a1, a3))
    assert s["p99"] <= 1e-5, s


def test_small_systems_fall_back(ctx):
    # below two superblocks (of 1024 bodies) the one-sided kernel runs
    pos, _, mass, h = galaxy_system(300, 700, seed=6)
    p32, m32, h32 = pos.astype(np.float32), mass.astype(np.float32), h.astype(np.float32)
    acc, _ = _run(ctx, p32, m32, h32, _sym(), phi=False)
    ref, _ = orc.accel(p32, h32, p32, m32, h32, want_phi=False)
    assert err_stats(rel_err(acc, ref))["p99"] <= 1e-5


def _uniform_mass_system(n, seed):
    # every mass 1/n, softening drawn per body: superblocks are flagged
    # uniform in mass but not in softening
    rng = np.random.default_rng(seed)
    p32 = rng.normal(size=(3, n)).astype(np.float32)
    m32 = np.full(n, 1.0 / n, np.float32)
    h32 = rng.uniform(5e-6, 5e-5, n).astype(np.float32)
    return p32, m32, h32


@pytest.mark.parametrize("n,tune", [(9000, "sym4"), (65536, "sym8"), (65536, None)])
def test_uniform_mass_mixed_softening(ctx, n, tune):
    # off-diagonal items of such superblocks take the generic pair loop, whose
    # sums already carry m_j (i side) and m_i (j side); a second mass factor
    # would scale every such contribution by 1/n
    p32, m32, h32 = _uniform_mass_system(n, seed=n)
    t = {"sym4": _sym(4), "sym8": _sym(8), None: None}[tune]
    # tune None: 65536 bodies make 528 items of 2048-body superblocks, so
    # khrr_accel takes the pair-symmetric path by itself
    acc, phi = _run(ctx, p32, m32, h32, t)
    rows = np.random.default_rng(1).choice(n, size=min(n, 4096), replace=False)
    ref_acc, ref_phi = orc.accel_rows(rows, p32, h32, m32)
    s = err_stats(rel_err(acc[:, rows], ref_acc))
    print("uniform mass, mixed softening", n, tune, s)
    assert s["median"] <= 1e-6 and s["p99"] <= 1e-5, s
    assert np.max(np.abs(phi[rows] - ref_phi) / np.abs(ref_phi)) <= 2e-6


def test_uniform_mass_mixed_softening_batched(ctx):
    # the same systems through the batched form: each equals its plain
    # pair-symmetric launch bit for bit and the oracle within the bar
    import torch
    import khrr
    sys_ = [_uniform_mass_system(n, seed=7 + n) for n in (20000, 40000, 3000)]
    off = np.concatenate([[0], np.cumsum([s[0].shape[1] for s in sys_])]).astype(np.int32)
    pos = np.concatenate([s[0] for s in sys_], axis=1)
    m = np.concatenate([s[1] for s in sys_])
    h = np.concatenate([s[2] for s in sys_])
    N = pos.shape[1]
    acc = torch.zeros((3, N), device="cuda:0")
    ph = torch.zeros(N, device="cuda:0")
    ctx.accel_batched(to_dev(off, torch.int32), off.tolist(), to_dev(pos, torch.float32),
                      to_dev(h, torch.float32), to_dev(m, torch.float32), acc=acc, phi=ph)
    torch.cuda.synchronize()
    acc, ph = acc.cpu().numpy(), ph.cpu().numpy()
    for k, (p32, m32, h32) in enumerate(sys_):
        a, b = off[k], off[k + 1]
        n = b - a
        plain = khrr.Tune(8, 4, 1, 256, 1) if n >= 16384 else \
            khrr.Tune(4, 4, max(1, min(32, -(-n // 1024))), 256, 0)
        pa, pp = _run(ctx, p32, m32, h32, plain)
        np.testing.assert_array_equal(acc[:, a:b], pa.astype(np.float32), err_msg=f"system {k}")
        np.testing.assert_array_equal(ph[a:b], pp.astype(np.float32), err_msg=f"system {k}")
        rows = np.random.default_rng(k).choice(n, size=min(n, 2048), replace=False)
        ref_acc, _ = orc.accel_rows(rows, p32, h32, m32)
        s = err_stats(rel_err(acc[:, a:b][:, rows], ref_acc))
        assert s["median"] <= 1e-6 and s["p99"] <= 1e-5, (k, s)


@pytest.mark.parametrize("phi", [False, True], ids=["acc", "acc+phi"])
def test_compensated_ipl8_superblocks(ctx, phi):
    # 8 i per lane over 2048-body superblocks: the compensated sums' TwoSum
    # error terms share the LDS with the j chunk and the j-side sums; both
    # forms within the bar, the compensated one at least as accurate
    import khrr
    pos, _, mass, h = galaxy_system(16000, 24000, seed=12)
    p32, m32, h32 = pos.astype(np.float32), mass.astype(np.float32), h.astype(np.float32)
    n = p32.shape[1]
    t = khrr.Tune(8, 4, 1, 256, 1, 2048)
    acc_c, phi_c = _run(ctx, p32, m32, h32, t, phi=phi, compensated=True)
    acc_p, _ = _run(ctx, p32, m32, h32, t, phi=phi)
    rows = np.sort(np.random.default_rng(3).choice(n, size=4096, replace=False))
    ref_acc, ref_phi = orc.accel_rows(rows, p32, h32, m32)
    sc = err_stats(rel_err(acc_c[:, rows], ref_acc))
    sp = err_stats(rel_err(acc_p[:, rows], ref_acc))
    print("ipl 8 / 2048: compensated", sc, "plain", sp)
    assert sc["median"] <= 1e-6 and sc["p99"] <= 1e-5, sc
    assert sp["median"] <= 1e-6 and sp["p99"] <= 1e-5, sp
    assert sc["median"] <= sp["median"] and sc["p99"] <= sp["p99"]
    if phi:
        assert np.max(np.abs(phi_c[rows] - ref_phi) / np.abs(ref_phi)) <= 2e-6
