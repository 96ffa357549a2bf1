"""Conservation over the small plain-method workload: energy, momentum and
angular-momentum drift of the GPU step loop against the FP64 oracle's run of
the same bodies and steps (tests/golden/galaxy24k_drift_golden.npz, written by
make_drift_golden.py).

Workload: one galaxy without external potential, 8192 disk stars + 16384
Plummer halo bodies, eps 0.004, dt 0.002, 500 predictor-corrector steps
(step order nbody.cu:1525-1582, force accumulator cleared every step),
diagnostics every 10 steps (result() nbody.cu:797-833, phi including the
self term, PHI_kernel nbody-kernel.cuh:43-86).

Tolerances:
  energy      max_t |E(t)/E(0) - 1| of the GPU run <= 1.25 x the oracle's
              own maximum + 2e-6 (the drift is the integrator's, set by dt;
              FP32 forces perturb the trajectory, not the scheme); E(0)
              within 1e-6 relative of the oracle's (FP32 potential)
  momentum    max_t |P(t) - P(0)| <= 1e-7 x sum m|v| (the oracle's FP64
              forces cancel pairwise to ~1e-15; FP32 pair forces leave a
              residual of their rounding, which the compensated sums shrink)
  ang. mom.   max_t |Lz(t)/Lz(0) - 1| <= 1e-7 and within 1e-7 of the
              oracle's own Lz drift
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import galaxy_system, to_dev

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(GOLDEN, "galaxy24k_drift_golden.npz")
NAMES = ("Ek", "Ep", "Px", "Py", "Pz", "Lx", "Ly", "Lz")


def _gpu_run(ctx, pos, vel, mass, h, dt, steps, every, compensated):
    import torch
    f64 = dict(device="cuda:0", dtype=torch.float64)
    f32 = dict(device="cuda:0", dtype=torch.float32)
    n = pos.shape[1]
    P, V = to_dev(pos, torch.float64), to_dev(vel, torch.float64)
    M64 = to_dev(mass, torch.float64)
    m32, h32 = to_dev(mass, torch.float32), to_dev(h, torch.float32)
    P32 = P.float().contiguous()
    acc, acct = torch.zeros((3, n), **f32), torch.zeros((3, n), **f32)
    post, velt = torch.zeros((3, n), **f64), torch.zeros((3, n), **f64)
    phi = torch.zeros(n, **f32)

    def grav(out=None, want_phi=False):
        ctx.accel(P32, h32, P32, m32, h32, acc=out, phi=phi if want_phi else None,
                  compensated=compensated)

    def diag():
        grav(want_phi=True)
        d = ctx.diag(P, V, M64, phi)
        return [d[k] for k in NAMES]

    grav(acc)
    rows = [diag()]
    t = 0.0
    for s in range(steps):
        ctx.integrate(0, dt, t, acc, post, velt, P, V, pos32=P32)
        grav(acct)
        ctx.integrate(1, dt, t + dt, acct, P, V, post, velt, acc0=acc)
        t += dt
        if (s + 1) % every == 0:
            rows.append(diag())
    torch.cuda.synchronize()
    return np.array(rows)


def _drifts(rows):
    E = 0.5 * (rows[:, 0] + rows[:, 1])
    dE = np.abs(E / E[0] - 1.0)
    dP = np.linalg.norm(rows[:, 2:5] - rows[0, 2:5], axis=1)
    dLz = np.abs(rows[:, 7] / rows[0, 7] - 1.0)
    return E, dE, dP, dLz


@pytest.mark.parametrize("compensated", [False, True], ids=["plain", "compensated"])
def test_galaxy24k_drift_within_oracle(ctx, compensated):
    g = np.load(FIXTURE)
    pos, vel, mass, h = galaxy_system(int(g["n_s"]), int(g["n_dm"]), seed=int(g["seed"]))
    m32 = mass.astype(np.float32).astype(np.float64)
    h32 = h.astype(np.float32).astype(np.float64)
    chk = np.array([pos.sum(), vel.sum(), (pos * pos).sum(), (vel * vel).sum(), m32.sum(),
                    h32.sum()])
    np.testing.assert_array_equal(chk, g["ic_checksum"])  # same bodies as the fixture
    steps, every, dt = int(g["steps"]), int(g["every"]), float(g["dt"])
    rows = _gpu_run(ctx, pos, vel, mass, h, dt, steps, every, compensated)
    ref = g["diag"][:, :8]
    assert rows.shape == ref.shape
    E, dE, dP, dLz = _drifts(rows)
    E_o, dE_o, dP_o, dLz_o = _drifts(ref)
    p_scale = float(np.sum(m32 * np.linalg.norm(vel, axis=0)))
    print(f"\n[{'compensated' if compensated else 'plain'}] max|dE/E| gpu {dE.max():.3e} "
          f"oracle {dE_o.max():.3e}; max|dP| gpu {dP.max():.3e} oracle {dP_o.max():.3e} "
          f"(scale {p_scale:.3e}); max|dLz/Lz| gpu {dLz.max():.3e} oracle {dLz_o.max():.3e}")
    assert abs(E[0] / E_o[0] - 1.0) <= 1e-6
    assert dE.max() <= 1.25 * dE_o.max() + 2e-6
    assert dP.max() <= 1e-7 * p_scale
    assert dLz.max() <= 1e-7
    assert abs(dLz.max() - dLz_o.max()) <= 1e-7
