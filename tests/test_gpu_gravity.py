"""GPU parity of khrr_accel (ACCEL_kernel / PHI_kernel,
nbody-kernel.cuh:43-142) against the FP64 oracle.

Tolerances (FP32 pair arithmetic vs FP64 on the same FP32 inputs):
  acceleration  per-particle |a - a64| / |a64|: median <= 1e-6, p99 <= 1e-5
  potential     per-particle |phi - phi64| / |phi64|: max <= 2e-6
"""
import numpy as np
import pytest

import oracle as orc
from helpers import err_stats, galaxy_system, rel_err, to_dev

pytestmark = pytest.mark.gpu

ACC_MEDIAN, ACC_P99, PHI_MAX = 1e-6, 1e-5, 2e-6


def _f32_inputs(pos, mass, h):
    """The FP32 values the GPU sees, and the same values widened."""
    p32 = pos.astype(np.float32)
    m32 = mass.astype(np.float32)
    h32 = h.astype(np.float32)
    return p32, m32, h32


def _gpu_accel(ctx, p32, m32, h32, rows=None, **kw):
    import torch
    pos = to_dev(p32, torch.float32)
    m = to_dev(m32, torch.float32)
    h = to_dev(h32, torch.float32)
    if rows is None:
        pos_i, h_i = pos, h
    else:
        pos_i = to_dev(p32[:, rows], torch.float32)
        h_i = to_dev(h32[rows], torch.float32)
    n_i = pos_i.shape[1]
    acc = torch.zeros((3, n_i), device="cuda:0", dtype=torch.float32)
    phi = torch.zeros(n_i, device="cuda:0", dtype=torch.float32)
    ctx.accel(pos_i, h_i, pos, m, h, acc=acc, phi=phi, **kw)
    torch.cuda.synchronize()
    return acc.cpu().numpy().astype(np.float64), phi.cpu().numpy().astype(np.float64)


def test_galaxy24k_all_rows(ctx):
    # the 8192 + 16384-body galaxy, every row checked
    pos, _, mass, h = galaxy_system(8192, 16384, seed=12345)
    p32, m32, h32 = _f32_inputs(pos, mass, h)
    acc, phi = _gpu_accel(ctx, p32, m32, h32)
    ref_acc, ref_phi = orc.accel(p32, h32, p32, m32, h32)
    s = err_stats(rel_err(acc, ref_acc))
    print("accel error", s)
    assert s["median"] <= ACC_MEDIAN and s["p99"] <= ACC_P99
    assert np.max(np.abs(phi - ref_phi) / np.abs(ref_phi)) <= PHI_MAX


def test_compensated_is_at_least_as_accurate(ctx):
    pos, _, mass, h = galaxy_system(4096, 8192, seed=7)
    p32, m32, h32 = _f32_inputs(pos, mass, h)
    ref_acc, _ = orc.accel(p32, h32, p32, m32, h32)
    plain, _ = _gpu_accel(ctx, p32, m32, h32)
    comp, _ = _gpu_accel(ctx, p32, m32, h32, compensated=True)
    e_plain = err_stats(rel_err(plain, ref_acc))
    e_comp = err_stats(rel_err(comp, ref_acc))
    print("plain", e_plain, "compensated", e_comp)
    assert e_comp["p99"] <= ACC_P99 and e_comp["median"] <= ACC_MEDIAN
    assert e_comp["p99"] <= e_plain["p99"] * 1.05


@pytest.mark.parametrize("ipl", [1, 2, 4, 8])
@pytest.mark.parametrize("split", [1, 3])
def test_launch_shapes_agree(ctx, ipl, split):
    import khrr
    # ragged sizes: neither n_i nor n_j is a multiple of 64 or of the tile
    rng = np.random.default_rng(ipl * 10 + split)
    n = 3001
    pos = rng.normal(size=(3, n)).astype(np.float32)
    m = rng.uniform(0.5, 1.5, n).astype(np.float32) / n
    h = np.where(rng.uniform(size=n) < 0.3, 8e-6, 2e-5).astype(np.float32)  # mixed softening
    tune = khrr.Tune(ipl, 2, split, 256)
    acc, phi = _gpu_accel(ctx, pos, m, h, tune=tune)
    ref_acc, ref_phi = orc.accel(pos, h, pos, m, h)
    s = err_stats(rel_err(acc, ref_acc))
    assert s["p99"] <= ACC_P99, s
    assert np.max(np.abs(phi - ref_phi) / np.abs(ref_phi)) <= PHI_MAX


def test_sampled_rows_subset_i(ctx):
    # i-set != j-set: the force-check path of the headline workload
    pos, _, mass, h = galaxy_system(16384, 32768, seed=3)
    p32, m32, h32 = _f32_inputs(pos, mass, h)
    rows = np.sort(np.random.default_rng(0).choice(p32.shape[1], 1000, replace=False))
    acc, phi = _gpu_accel(ctx, p32, m32, h32, rows=rows)
    ref_acc, ref_phi = orc.accel_rows(rows, p32, h32, m32)
    s = err_stats(rel_err(acc, ref_acc))
    assert s["median"] <= ACC_MEDIAN and s["p99"] <= ACC_P99, s


def test_accumulate_and_overwrite(ctx):
    import torch
    rng = np.random.default_rng(11)
    n = 777
    p = rng.normal(size=(3, n)).astype(np.float32)
    m = np.full(n, 1.0 / n, np.float32)
    h = np.full(n, 1e-4, np.float32)
    pos, mm, hh = (to_dev(v, torch.float32) for v in (p, m, h))
    a0 = rng.normal(size=(3, n)).astype(np.float32)
    f0 = rng.normal(size=n).astype(np.float32)
    acc = to_dev(a0, torch.float32)
    phi = to_dev(f0, torch.float32)
    ctx.accel(pos, hh, pos, mm, hh, acc=acc, phi=phi, accumulate=True)
    acc2 = torch.zeros_like(acc)
    phi2 = torch.zeros_like(phi)
    ctx.accel(pos, hh, pos, mm, hh, acc=acc2, phi=phi2)
    torch.cuda.synchronize()
    # ACCEL adds (nbody-kernel.cuh:137-141), PHI subtracts (:85)
    np.testing.assert_array_equal(acc.cpu().numpy(), a0 + acc2.cpu().numpy())
    np.testing.assert_array_equal(phi.cpu().numpy(), f0 + phi2.cpu().numpy())


def test_phi_only_and_acc_only_match_fused(ctx):
    import torch
    rng = np.random.default_rng(12)
    n = 2048
    p = rng.normal(size=(3, n)).astype(np.float32)
    m = np.full(n, 1.0 / n, np.float32)
    h = np.full(n, 1e-4, np.float32)
    pos, mm, hh = (to_dev(v, torch.float32) for v in (p, m, h))
    a1 = torch.zeros((3, n), device="cuda:0")
    f1 = torch.zeros(n, device="cuda:0")
    ctx.accel(pos, hh, pos, mm, hh, acc=a1, phi=f1)
    a2 = torch.zeros((3, n), device="cuda:0")
    f2 = torch.zeros(n, device="cuda:0")
    ctx.accel(pos, hh, pos, mm, hh, acc=a2)
    ctx.accel(pos, hh, pos, mm, hh, phi=f2)
    torch.cuda.synchronize()
    assert torch.equal(a1, a2) and torch.equal(f1, f2)


def test_empty_sets(ctx):
    import torch
    z = torch.zeros((3, 0), device="cuda:0")
    z1 = torch.zeros(0, device="cuda:0")
    pos = torch.ones((3, 5), device="cuda:0")
    h = torch.full((5,), 1e-3, device="cuda:0")
    acc = torch.full((3, 5), 7.0, device="cuda:0")
    phi = torch.full((5,), 7.0, device="cuda:0")
    ctx.accel(pos, h, z, z1, z1, acc=acc, phi=phi, accumulate=True)  # no j: unchanged
    torch.cuda.synchronize()
    assert torch.all(acc == 7.0) and torch.all(phi == 7.0)
    ctx.accel(pos, h, z, z1, z1, acc=acc, phi=phi)  # overwrite: empty sum = 0
    torch.cuda.synchronize()
    assert torch.all(acc == 0.0) and torch.all(phi == 0.0)
    ctx.accel(z, z1, pos, h, h, acc=torch.zeros((3, 0), device="cuda:0"))  # no i


def test_deterministic(ctx):
    pos, _, mass, h = galaxy_system(2048, 4096, seed=5)
    p32, m32, h32 = _f32_inputs(pos, mass, h)
    a1, f1 = _gpu_accel(ctx, p32, m32, h32)
    a2, f2 = _gpu_accel(ctx, p32, m32, h32)
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(f1, f2)


def test_golden_fixture(ctx):
    import os
    from conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "gravity_golden.npz"))
    p32, m32, h32 = (g[k].astype(np.float32) for k in ("pos", "m", "h"))
    acc, phi = _gpu_accel(ctx, p32, m32, h32)
    # golden values are for the FP64 inputs; FP32 rounding of the inputs
    # adds ~1e-7 relative
    s = err_stats(rel_err(acc, g["acc"]))
    assert s["p99"] <= 2e-5, s
    assert np.max(np.abs(phi - g["phi"]) / np.abs(g["phi"])) <= 5e-6


def test_errors_are_reported(ctx):
    import khrr
    import torch
    pos = torch.zeros((3, 4), device="cuda:0")
    h = torch.zeros(4, device="cuda:0")
    with pytest.raises(khrr.KhrrError, match="both NULL"):
        ctx.accel(pos, h, pos, h, h)
    with pytest.raises(khrr.KhrrError, match="i_per_lane"):
        ctx.accel(pos, h, pos, h, h, acc=torch.zeros((3, 4), device="cuda:0"),
                  tune=khrr.Tune(3, 1, 1, 256))
