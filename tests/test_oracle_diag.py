"""Pins the oracle's diagnostics (result(), nbody.cu:797-833) on hand-made
states with closed-form values."""
import numpy as np
import pytest

import oracle as orc


def test_two_particles_closed_form():
    pos = np.array([[1.0, -1.0], [0.0, 0.0], [0.5, -0.5]])
    vel = np.array([[0.0, 0.0], [1.0, -1.0], [0.0, 0.0]])
    m = np.array([2.0, 2.0])
    phi = np.array([-3.0, -3.0])
    d = orc.diag(pos, vel, m, phi)
    Ek, Ep, Px, Py, Pz, Lx, Ly, Lz, M, Mx, My, Mz, vr, vfi, vz, R, Z = d
    assert Ek == 4.0 and Ep == -12.0
    assert (Px, Py, Pz) == (0.0, 0.0, 0.0)
    # L = sum m r x v: r x v = (y vz - z vy, z vx - x vz, x vy - y vx)
    L = sum(m[i] * np.cross(pos[:, i], vel[:, i]) for i in range(2))
    np.testing.assert_allclose([Lx, Ly, Lz], L)
    assert (M, Mx, My, Mz) == (4.0, 0.0, 0.0, 0.0)
    assert vfi == 1.0 and vr == 0.0 and R == 1.0 and Z == 0.5 and vz == 0.0


def test_vr_vfi_signs():
    pos = np.array([[0.0], [2.0], [0.0]])
    vel = np.array([[-1.0], [3.0], [0.0]])
    d = orc.diag(pos, vel, np.ones(1))
    # vr = (vx x + vy y)/r = 3, vfi = (vy x - vx y)/r = 1
    assert d[12] == 3.0 and d[13] == pytest.approx(1.0)
    # particle on the axis: vr = vfi = 0
    d0 = orc.diag(np.zeros((3, 1)), np.ones((3, 1)), np.ones(1))
    assert d0[12] == 0.0 and d0[13] == 0.0


def test_sums_match_numpy():
    rng = np.random.default_rng(0)
    n = 1000
    pos, vel = rng.normal(size=(3, n)), rng.normal(size=(3, n))
    m = rng.uniform(size=n)
    phi = rng.normal(size=n)
    d = orc.diag(pos, vel, m, phi)
    np.testing.assert_allclose(d[0], (m * (vel ** 2).sum(0)).sum(), rtol=1e-13)
    np.testing.assert_allclose(d[1], (m * phi).sum(), rtol=1e-12)
    np.testing.assert_allclose(d[2:5], (m * vel).sum(1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(d[5:8], (m * np.cross(pos.T, vel.T).T).sum(1), rtol=1e-11,
                               atol=1e-12)
    np.testing.assert_allclose(d[9:12], (m * pos).sum(1), rtol=1e-12, atol=1e-12)
