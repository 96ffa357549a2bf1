"""Pins the oracle's gravity (ACCEL_kernel nbody-kernel.cuh:89-142, PHI_kernel
:43-86) against closed forms, symmetry, conservation and an independent
extended-precision restatement.  CPU only."""
import numpy as np
import pytest

import oracle as orc


def test_two_body_closed_form():
    # a_i = m_j d / (|d|^2 + h_i + h_j)^(3/2);  phi_i = -sum_j m_j / sqrt(.)
    pos = np.array([[0.1, 0.7], [-0.2, 0.4], [0.3, -0.5]])
    m = np.array([2.0, 3.0])
    h = np.array([1e-4, 4e-4])
    acc, phi = orc.accel(pos, h, pos, m, h)
    d = pos[:, 1] - pos[:, 0]
    r2 = d @ d + h[0] + h[1]
    np.testing.assert_allclose(acc[:, 0], m[1] * d / r2 ** 1.5, rtol=1e-14)
    np.testing.assert_allclose(acc[:, 1], -m[0] * d / r2 ** 1.5, rtol=1e-14)
    # self term m_i / sqrt(2 h_i) is included (PHI_kernel sums j = i)
    self0 = m[0] / np.sqrt(2 * h[0])
    np.testing.assert_allclose(phi[0], -(self0 + m[1] / np.sqrt(r2)), rtol=1e-14)


def test_self_pair_gives_no_force():
    pos = np.array([[0.5], [0.25], [-1.0]])
    acc, phi = orc.accel(pos, np.array([2e-3]), pos, np.array([7.0]), np.array([2e-3]))
    assert np.all(acc == 0.0)
    assert phi[0] == pytest.approx(-7.0 / np.sqrt(4e-3), rel=1e-15)


@pytest.mark.parametrize("k", [3, 5, 12])
def test_regular_polygon_centre_has_zero_force(k):
    # k equal masses on a circle around a test particle at the centre
    ang = 2 * np.pi * np.arange(k) / k
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(k)])
    pos = np.concatenate([np.zeros((3, 1)), ring], axis=1)
    m = np.concatenate([[0.0], np.ones(k)])
    h = np.full(k + 1, 1e-3)
    acc, _ = orc.accel(pos[:, :1], h[:1], pos, m, h)
    assert np.abs(acc).max() < 1e-14


def test_momentum_balance():
    # symmetric softening -> antisymmetric pair law -> sum_i m_i a_i = 0
    rng = np.random.default_rng(1)
    n = 300
    pos = rng.normal(size=(3, n))
    m = rng.uniform(0.5, 2.0, n)
    h = rng.uniform(1e-5, 1e-3, n)
    acc, _ = orc.accel(pos, h, pos, m, h)
    tot = (acc * m).sum(axis=1)
    scale = (np.abs(acc) * m).sum()
    assert np.abs(tot).max() < 1e-13 * scale


def test_far_field_monopole():
    rng = np.random.default_rng(2)
    n = 200
    pos = rng.uniform(-0.01, 0.01, (3, n))
    m = np.full(n, 1.0 / n)
    h = np.full(n, 1e-8)
    probe = np.array([[100.0], [0.0], [0.0]])
    acc, phi = orc.accel(probe, np.array([1e-8]), pos, m, h)
    assert acc[0, 0] == pytest.approx(-1.0 / 100.0 ** 2, rel=1e-3)
    assert phi[0] == pytest.approx(-1.0 / 100.0, rel=1e-3)


def test_translation_invariance():
    rng = np.random.default_rng(3)
    pos = rng.normal(size=(3, 64))
    m = rng.uniform(0.1, 1.0, 64)
    h = np.full(64, 1e-4)
    a1, p1 = orc.accel(pos, h, pos, m, h)
    a2, p2 = orc.accel(pos + np.array([[0.5], [-0.25], [0.125]]), h,
                       pos + np.array([[0.5], [-0.25], [0.125]]), m, h)
    np.testing.assert_allclose(a1, a2, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(p1, p2, rtol=1e-12)


def test_matches_independent_longdouble_restatement():
    # numpy extended-precision restatement of the same sum (pairwise, not
    # the oracle's loop) on random bodies
    rng = np.random.default_rng(4)
    n = 257
    pos = rng.normal(size=(3, n))
    m = rng.uniform(0.1, 1.0, n)
    h = rng.uniform(1e-5, 1e-3, n)
    acc, phi = orc.accel(pos, h, pos, m, h)
    P = pos.astype(np.longdouble)
    d = P[:, None, :] - P[:, :, None]  # d[c, i, j] = x_j - x_i
    r2 = (d * d).sum(axis=0) + h[:, None] + h[None, :]
    inv = 1 / np.sqrt(r2)
    ref_acc = (d * (m * inv ** 3)[None]).sum(axis=2)
    ref_phi = -(m * inv).sum(axis=1)
    np.testing.assert_allclose(acc, ref_acc.astype(np.float64), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(phi, ref_phi.astype(np.float64), rtol=1e-13)


def test_accumulate_semantics():
    # ACCEL_kernel adds into acc (:137-141); PHI_kernel subtracts (:85)
    rng = np.random.default_rng(5)
    pos = rng.normal(size=(3, 10))
    m = np.ones(10)
    h = np.full(10, 1e-3)
    a, p = orc.accel(pos, h, pos, m, h)
    a0 = rng.normal(size=(3, 10))
    p0 = rng.normal(size=10)
    a2, p2 = orc.accel(pos, h, pos, m, h, acc=a0, phi=p0, accumulate=True)
    np.testing.assert_allclose(a2, a0 + a, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(p2, p0 + p, rtol=1e-15)


def test_rows_match_full():
    rng = np.random.default_rng(6)
    pos = rng.normal(size=(3, 100))
    m = rng.uniform(size=100)
    h = np.full(100, 1e-4)
    a, p = orc.accel(pos, h, pos, m, h)
    rows = np.array([0, 17, 99, 42])
    ar, pr = orc.accel_rows(rows, pos, h, m)
    np.testing.assert_array_equal(ar, a[:, rows])
    np.testing.assert_array_equal(pr, p[rows])


def test_golden_vectors():
    import os
    from conftest import GOLDEN
    path = os.path.join(GOLDEN, "gravity_golden.npz")
    g = np.load(path)
    acc, phi = orc.accel(g["pos"], g["h"], g["pos"], g["m"], g["h"])
    np.testing.assert_allclose(acc, g["acc"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(phi, g["phi"], rtol=1e-13)
