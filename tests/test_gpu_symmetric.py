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
