"""Pins the oracle's stencil (wave_diss_iteration wave-kernel.cuh:5-103,
rotation nbody.cu:227-239) with invariants and a hand-evaluated cell."""
import numpy as np

import oracle as orc

PI = 3.14159265358979  # common.cuh:7


def _w(nx=10, dx=0.2):
    return orc.reference_wave(nx, dx)


def test_zero_in_zero_out():
    w = _w()
    G = w["nx"] ** 3
    out = orc.wave_step(w, np.zeros(G), np.zeros(G), np.zeros(G))
    assert np.all(out == 0)


def test_faces_are_zero_and_linear():
    w = _w(12)
    rng = np.random.default_rng(0)
    G = w["nx"] ** 3
    a = [rng.normal(size=G) for _ in range(3)]
    b = [rng.normal(size=G) for _ in range(3)]
    fa = orc.wave_step(w, *a)
    fb = orc.wave_step(w, *b)
    fab = orc.wave_step(w, *(x + y for x, y in zip(a, b)))
    np.testing.assert_allclose(fab, fa + fb, rtol=1e-12, atol=1e-9)
    f = fa.reshape(12, 12, 12)
    for face in (f[0], f[-1], f[:, 0], f[:, -1], f[:, :, 0], f[:, :, -1]):
        assert np.all(face == 0)


def test_one_cell_by_hand():
    # evaluate the update of one interior cell from the formula
    w = _w(8, 0.25)
    nx = w["nx"]
    G = nx ** 3
    rng = np.random.default_rng(1)
    phi = rng.normal(size=G)
    old = rng.normal(size=G)
    mass = rng.uniform(size=G)
    out = orc.wave_step(w, phi, old, mass)
    ix, iy, iz = 3, 6, 5
    at = lambda x, y, z: x + y * nx + z * nx * nx  # noqa: E731
    c, dt, dx = w["c"], w["dt"], w["dx"]
    edge = w["sim_l"] - w["bc_l"]
    x, y, z = ix * dx, iy * dx, iz * dx
    diss = w["diss_base"] + w["diss_extra"] * (x - edge) ** 2  # x > -edge always
    diss += w["diss_extra"] * (y - edge) ** 2 if y > edge else 0.0
    diss += w["diss_extra"] * (z - edge) ** 2 if z > edge else 0.0
    Q = 0.5 * diss * c * dt
    K = c * c * dt * dt / (dx * dx)
    nb = sum(phi[at(ix + a, iy + b, iz + cc)] for a, b, cc in
             ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)))
    f = 4 * PI * mass[at(ix, iy, iz)] / dx ** 3
    expect = (K * nb + 2 * (1 - 3 * K) * phi[at(ix, iy, iz)] + (Q - 1) * old[at(ix, iy, iz)]
              - c * c * dt * dt * f) / (1 + Q)
    assert out[at(ix, iy, iz)] == np.float64(expect) or abs(out[at(ix, iy, iz)] - expect) <= \
        1e-13 * abs(expect)
    # CFL of the reference constants: K = 1/12
    assert abs(K - 1.0 / 12.0) < 1e-15


def test_iterate_rotation_matches_manual_steps():
    w = _w(10)
    G = w["nx"] ** 3
    rng = np.random.default_rng(2)
    mass = np.zeros(G)
    mass[G // 2 + 55] = 1.0
    prev, curr = np.zeros(G), rng.normal(size=G) * 1e-3
    out, prev_out = orc.wave_iterate(w, prev, curr, mass, 4)
    p, c = prev, curr
    for _ in range(4):
        n = orc.wave_step(w, c, p, mass)
        p, c = c, n
    np.testing.assert_array_equal(out, c)
    np.testing.assert_array_equal(prev_out, p)


def test_point_mass_builds_negative_potential_well():
    # with a source the field goes negative at the mass (attractive
    # potential) and stays bounded (damped, CFL-stable)
    w = _w(16, 0.25)
    nx = w["nx"]
    G = nx ** 3
    mass = np.zeros(G)
    centre = 8 + 8 * nx + 8 * nx * nx
    mass[centre] = 1e-3
    out, _ = orc.wave_iterate(w, np.zeros(G), np.zeros(G), mass, 300)
    assert out[centre] < 0 and out[centre] == out.min()
    assert np.all(np.isfinite(out))
