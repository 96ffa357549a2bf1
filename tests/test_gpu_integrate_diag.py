"""GPU parity of khrr_integrate (kernelNbody_integTime, nbody-kernel.cuh:144-194)
and khrr_diag (result() sums, nbody.cu:797-833) against the FP64 oracle.

Tolerances:
  integrator without external potential: bit-exact (same FP64 operations,
    contraction off on both sides)
  integrator with halo + bulge (dev_fex uses the device atan/log): 1e-13
    relative per component
  diagnostics: sums 1e-12 relative to the sum of magnitudes (FP64
    reassociation only), maxima exact
"""
import numpy as np
import pytest

import oracle as orc
from helpers import galaxy_system, to_dev

pytestmark = pytest.mark.gpu


def _state(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(3, n))
    vel = rng.normal(size=(3, n))
    acc = rng.normal(size=(3, n)).astype(np.float32)
    return pos, vel, acc


@pytest.mark.parametrize("istep", [0, 1])
def test_integrator_bit_exact_no_external(ctx, istep):
    import torch
    n = 10007
    pos, vel, acc = _state(n, istep + 1)
    pos_t0, vel_t0, _ = _state(n, istep + 50)
    dt = 0.002
    d = {k: to_dev(v, torch.float64) for k, v in
         dict(pos=pos, vel=vel, pos_t=pos_t0, vel_t=vel_t0).items()}
    accd = to_dev(acc, torch.float32)
    acc0 = torch.zeros((3, n), device="cuda:0")
    pos32 = torch.zeros((3, n), device="cuda:0")
    ctx.integrate(istep, dt, 0.0, accd, d["pos_t"], d["vel_t"], d["pos"], d["vel"], acc0=acc0,
                  pos32=pos32)
    torch.cuda.synchronize()
    r_pos_t, r_vel_t, r_acc0 = pos_t0.copy(), vel_t0.copy(), np.zeros((3, n))
    orc.integrate(istep, dt, 0.0, acc.astype(np.float64), r_pos_t, r_vel_t, pos, vel, r_acc0)
    np.testing.assert_array_equal(d["pos_t"].cpu().numpy(), r_pos_t)
    np.testing.assert_array_equal(d["vel_t"].cpu().numpy(), r_vel_t)
    np.testing.assert_array_equal(pos32.cpu().numpy(), r_pos_t.astype(np.float32))
    if istep == 1:
        np.testing.assert_array_equal(acc0.cpu().numpy(), acc)


def test_integrator_with_halo_and_bulge(ctx):
    import khrr
    import torch
    n = 4096
    pos, vel, acc = _state(n, 9)
    pos *= 0.5  # both sides of Rb = 0.2 and Rh2 = 3
    pos[:, :100] *= 20.0
    ext = khrr.extpot_init(2.0, 0.3333, 1.0, 0.5, 0.0431, 0.2)
    oext = orc.extpot(2.0, 0.3333, 1.0, 0.5, 0.0431, 0.2)
    d = {k: to_dev(v, torch.float64) for k, v in dict(pos=pos, vel=vel).items()}
    pos_t = torch.zeros((3, n), device="cuda:0", dtype=torch.float64)
    vel_t = torch.zeros((3, n), device="cuda:0", dtype=torch.float64)
    ctx.integrate(0, 0.002, 0.0, to_dev(acc, torch.float32), pos_t, vel_t, d["pos"], d["vel"],
                  ext=ext)
    torch.cuda.synchronize()
    r_pos_t, r_vel_t = np.zeros((3, n)), np.zeros((3, n))
    orc.integrate(0, 0.002, 0.0, acc.astype(np.float64), r_pos_t, r_vel_t, pos, vel, ext=oext)
    np.testing.assert_allclose(vel_t.cpu().numpy(), r_vel_t, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(pos_t.cpu().numpy(), r_pos_t, rtol=1e-13, atol=1e-15)


def _check_diag(out, pos, vel, mass, phi, tol=1e-13):
    import khrr
    ref = orc.diag(pos, vel, mass, phi.astype(np.float64))
    # scale of each sum: sum of |terms| (products taken term by term, so the
    # cross-product components do not cancel in the scale)
    x, y, z = np.abs(pos)
    vx, vy, vz = np.abs(vel)
    scale = [np.sum(mass * (vx * vx + vy * vy + vz * vz)), np.sum(mass * np.abs(phi)),
             np.sum(mass * vx), np.sum(mass * vy), np.sum(mass * vz),
             np.sum(mass * (vz * y + vy * z)), np.sum(mass * (vx * z + vz * x)),
             np.sum(mass * (vy * x + vx * y)), np.sum(mass), np.sum(mass * x),
             np.sum(mass * y), np.sum(mass * z)]
    for k, name in enumerate(khrr.DIAG_NAMES):
        if k < 12:
            assert abs(out[name] - ref[k]) <= tol * scale[k], name
        else:
            assert out[name] == ref[k], name


def test_diag_matches_oracle(ctx):
    import torch
    pos, vel, mass, _ = galaxy_system(3000, 5000, seed=21)
    phi = -np.random.default_rng(1).uniform(1, 10, pos.shape[1]).astype(np.float32)
    out = ctx.diag(to_dev(pos, torch.float64), to_dev(vel, torch.float64),
                   to_dev(mass, torch.float64), to_dev(phi, torch.float32))
    _check_diag(out, pos, vel, mass, phi)


def test_diag_large_ragged(ctx):
    # Above 1024 * 1024 bodies the slice count stops at 1024, so each thread
    # walks its slice in more than one group of four bodies and the last
    # group is cut short; 1300001 bodies also leave the last slice shorter.
    # The oracle's serial FP64 sums over this many terms carry rounding of
    # about 2e-14 of the sum of magnitudes, so the bar is 1e-12.
    import torch
    n = 1300001
    rng = np.random.default_rng(1300)
    pos = rng.normal(size=(3, n))
    vel = rng.normal(size=(3, n))
    mass = rng.uniform(0.5, 1.5, n)
    phi = -rng.uniform(1, 10, n).astype(np.float32)
    out = ctx.diag(to_dev(pos, torch.float64), to_dev(vel, torch.float64),
                   to_dev(mass, torch.float64), to_dev(phi, torch.float32))
    _check_diag(out, pos, vel, mass, phi, tol=1e-12)


def test_diag_empty_and_small(ctx):
    import torch
    z3 = torch.zeros((3, 0), device="cuda:0", dtype=torch.float64)
    z = torch.zeros(0, device="cuda:0", dtype=torch.float64)
    out = ctx.diag(z3, z3, z)
    assert all(v == 0.0 for v in out.values())
    pos = np.array([[1.0], [2.0], [3.0]])
    vel = np.array([[0.5], [-1.0], [2.0]])
    m = np.array([3.0])
    out = ctx.diag(to_dev(pos, torch.float64), to_dev(vel, torch.float64),
                   to_dev(m, torch.float64))
    # L = m r x v
    L = 3.0 * np.cross(pos[:, 0], vel[:, 0])
    assert out["Lx"] == pytest.approx(L[0]) and out["Lz"] == pytest.approx(L[2])
    assert out["Ek"] == pytest.approx(3.0 * (0.25 + 1 + 4))
    assert out["M"] == 3.0 and out["Mz"] == 9.0
