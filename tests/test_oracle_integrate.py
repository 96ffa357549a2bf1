"""Pins the oracle's integrator (kernelNbody_integTime nbody-kernel.cuh:144-194,
dev_fex :8-40, constants nbody.cu:1111-1120, loop nbody.cu:1468-1594)."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN


def test_constant_force_is_integrated_exactly():
    # f constant: predictor gives x + v dt + f dt^2/2, corrector v + f dt
    rng = np.random.default_rng(0)
    n = 50
    x = rng.normal(size=(3, n))
    v = rng.normal(size=(3, n))
    f = rng.normal(size=(3, n))
    dt = 0.25  # exact binary fraction: no rounding in dt/2, dt^2
    xt, vt = np.zeros((3, n)), np.zeros((3, n))
    orc.integrate(0, dt, 0.0, f, xt, vt, x, v)
    np.testing.assert_allclose(xt, x + v * dt + 0.5 * f * dt * dt, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(vt, v + f * dt, rtol=1e-15, atol=1e-15)
    pos, vel, acc0 = np.zeros((3, n)), v.copy(), np.zeros((3, n))
    orc.integrate(1, dt, dt, f, pos, vel, xt, vt, acc0)
    np.testing.assert_allclose(vel, v + f * dt, rtol=1e-14, atol=1e-14)
    np.testing.assert_array_equal(pos, xt)
    np.testing.assert_array_equal(acc0, f)


def test_free_motion():
    x = np.array([[1.0], [2.0], [3.0]])
    v = np.array([[0.5], [-0.25], [0.125]])
    f = np.zeros((3, 1))
    xt, vt = np.zeros((3, 1)), np.zeros((3, 1))
    orc.integrate(0, 0.5, 0.0, f, xt, vt, x, v)
    np.testing.assert_array_equal(xt, x + 0.5 * v)
    np.testing.assert_array_equal(vt, v)


def test_kepler_orbit_against_mpmath_golden():
    # the full reference loop on a 2-body circular orbit vs a 40-digit
    # evaluation of the same recurrence (tests/golden/make_golden.py)
    g = np.load(os.path.join(GOLDEN, "kepler_golden.npz"))
    pos = np.array([[-0.5, 0.5], [0.0, 0.0], [0.0, 0.0]])
    vel = np.array([[0.0, 0.0], [-0.5, 0.5], [0.0, 0.0]])
    m = np.array([0.5, 0.5])
    h = np.full(2, 2.0 ** -34)
    p, v, d = orc.run(pos, vel, m, h, float(g["dt"]), int(g["steps"]))
    np.testing.assert_allclose(p, g["pos"], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(v, g["vel"], rtol=1e-11, atol=1e-12)
    # energy of the circular orbit: E = (Ek + Ep)/2 with the reference's
    # conventions; drift over 200 steps of dt = 0.01 stays tiny
    E = 0.5 * (d[:, 0] + d[:, 1])
    assert np.max(np.abs(E / E[0] - 1)) < 1e-6
    # momentum exactly conserved by the symmetric pair law
    assert np.max(np.abs(d[:, 2:5])) < 1e-15


def test_galaxy24k_drift_fixture_is_the_oracle_run():
    # the committed 500-step drift fixture (make_drift_golden.py) holds the
    # oracle's own run of these bodies: same initial conditions, and its
    # first two rows (t = 0 and after 10 steps) equal a fresh run bit for bit
    g = np.load(os.path.join(GOLDEN, "galaxy24k_drift_golden.npz"))
    from helpers import galaxy_system
    pos, vel, mass, h = galaxy_system(int(g["n_s"]), int(g["n_dm"]), seed=int(g["seed"]))
    m32 = mass.astype(np.float32).astype(np.float64)
    h32 = h.astype(np.float32).astype(np.float64)
    chk = np.array([pos.sum(), vel.sum(), (pos * pos).sum(), (vel * vel).sum(), m32.sum(),
                    h32.sum()])
    np.testing.assert_array_equal(chk, g["ic_checksum"])
    every = int(g["every"])
    _, _, d = orc.run(pos, vel, m32, h32, float(g["dt"]), every, diag_every=every)
    np.testing.assert_array_equal(d, g["diag"][:2])
    # the reference method conserves what it should over the 500 steps:
    # energy to the integrator's error, momentum to FP64 rounding
    E = 0.5 * (g["diag"][:, 0] + g["diag"][:, 1])
    assert np.max(np.abs(E / E[0] - 1)) < 1e-3
    p_scale = float(np.sum(m32 * np.linalg.norm(vel, axis=0)))
    assert np.max(np.abs(g["diag"][:, 2:5] - g["diag"][0, 2:5])) < 1e-12 * p_scale


def test_accumulator_reset_vs_compat_mode():
    # the never-reset force accumulator of nbody.cu:1560 blows up energy
    rng = np.random.default_rng(1)
    import ics
    pos, vel = ics.plummer(256, 1.0, 0.3333, 1.0, seed=3)
    m = np.full(256, 1.0 / 256)
    h = np.full(256, 0.5 * 0.01 ** 2)
    _, _, d_ok = orc.run(pos, vel, m, h, 0.002, 20)
    _, _, d_bad = orc.run(pos, vel, m, h, 0.002, 20, compat_accumulate=True)
    E_ok = 0.5 * (d_ok[:, 0] + d_ok[:, 1])
    E_bad = 0.5 * (d_bad[:, 0] + d_bad[:, 1])
    assert abs(E_ok[-1] / E_ok[0] - 1) < 1e-4
    assert abs(E_bad[-1] / E_bad[0] - 1) > 1e-2
    del rng


@pytest.fixture
def ext():
    return orc.extpot(2.0, 0.3333, 1.0, 0.5, 0.0431, 0.2)


def test_dev_fex_zero_without_halo_and_bulge():
    e = orc.extpot(0.0, 0.3333, 1.0, 0.0, 0.0431, 0.2)
    for p in ([0.1, 0.2, 0.3], [5.0, 0.0, 0.0], [0.0, 0.0, 0.0]):
        assert np.all(orc.dev_fex(e, p) == 0.0)


def test_dev_fex_continuous_at_branch_radii(ext):
    # halo branch switches at Rh2 = 3 Rh, bulge at Rb; Mh_inf and const1
    # are defined so that both sides agree (nbody.cu:1116-1120)
    for R in (ext.Rh2, ext.Rb):
        d = np.array([1.0, 2.0, 2.0]) / 3.0
        fin = orc.dev_fex(ext, d * R * (1 - 1e-12))
        fout = orc.dev_fex(ext, d * R * (1 + 1e-12))
        np.testing.assert_allclose(fin, fout, rtol=1e-9)


def test_dev_fex_point_mass_outside(ext):
    # outside both radii the force is -(Mh_inf + Mb) r / r^3
    p = np.array([10.0, -4.0, 2.0])
    r = np.linalg.norm(p)
    np.testing.assert_allclose(orc.dev_fex(ext, p), -(ext.Mh_inf + ext.Mb) * p / r ** 3,
                               rtol=1e-14)
    assert ext.Mh_inf > ext.Mh  # the halo keeps growing out to Rh2


def test_dev_fex_halo_closed_form(ext):
    # inside Rh2 the halo term is -con (x - atan x)/r^3 with x = r/a; the
    # enclosed mass at Rh is Mh by construction of con
    r = 1.0  # = Rh
    p = np.array([r, 0.0, 0.0])
    e_halo_only = orc.extpot(2.0, 0.3333, 1.0, 0.0, 0.0431, 0.2)
    f = orc.dev_fex(e_halo_only, p)
    assert f[0] == pytest.approx(-2.0 / r ** 2, rel=1e-13)
