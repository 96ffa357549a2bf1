"""Pins of the oracle's mesh force (acc_field, p2mesh-kernel.cuh:169-224,
evaluated on the full 3-D grid; apply_acceleration_to_particles /
apply_cell_to_particles, :226-252) against closed forms and numpy."""
import numpy as np
import pytest

import oracle as orc


def _coords(nx, dx):
    i = np.arange(nx) * dx
    z, y, x = np.meshgrid(i, i, i, indexing="ij")  # x fastest in memory
    return x.reshape(-1), y.reshape(-1), z.reshape(-1)


def test_linear_field_gives_constant_acceleration():
    nx, dx = 11, 0.37
    x, y, z = _coords(nx, dx)
    phi = 1.5 * x - 0.25 * y + 3.0 * z + 7.0
    acc, ab = orc.mesh_acc(nx, dx, phi)
    np.testing.assert_allclose(acc[0], -1.5, rtol=1e-12)
    np.testing.assert_allclose(acc[1], 0.25, rtol=1e-12)
    np.testing.assert_allclose(acc[2], -3.0, rtol=1e-12)
    np.testing.assert_allclose(ab, np.sqrt(1.5 ** 2 + 0.25 ** 2 + 9.0), rtol=1e-12)


def test_quadratic_field_central_inside_one_sided_on_faces():
    nx, dx = 9, 0.5
    x, y, z = _coords(nx, dx)
    acc, _ = orc.mesh_acc(nx, dx, x * x, want_abs=False)
    ix = np.rint(x / dx).astype(int)
    want = -2.0 * x  # central difference of x^2 is exact
    want = np.where(ix == 0, -(2.0 * x + dx), want)       # forward difference
    want = np.where(ix == nx - 1, -(2.0 * x - dx), want)  # backward difference
    np.testing.assert_allclose(acc[0], want, rtol=1e-13, atol=1e-13)
    assert np.all(acc[1] == 0.0) and np.all(acc[2] == 0.0)


def test_matches_numpy_gradient():
    nx, dx = 13, 0.2
    phi = np.random.default_rng(3).normal(size=nx ** 3)
    acc, ab = orc.mesh_acc(nx, dx, phi)
    gz, gy, gx = np.gradient(phi.reshape(nx, nx, nx), dx, edge_order=1)
    for k, g in enumerate((gx, gy, gz)):
        np.testing.assert_allclose(acc[k], -g.reshape(-1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ab, np.linalg.norm(acc, axis=0), rtol=1e-14)


def test_mesh_to_particles_is_cell_lookup():
    rng = np.random.default_rng(4)
    G = 125
    cell_acc = rng.normal(size=(3, G))
    cell_phi = rng.normal(size=G)
    cell_id = rng.integers(0, G, 1000).astype(np.int32)
    acc, phi = orc.mesh_to_particles(cell_id, cell_acc, cell_phi)
    np.testing.assert_array_equal(acc, cell_acc[:, cell_id])
    np.testing.assert_array_equal(phi, cell_phi[cell_id])


def test_a_wrong_sign_would_fail():
    # a plausible mistake (gradient instead of minus gradient) is caught
    nx, dx = 5, 1.0
    x, _, _ = _coords(nx, dx)
    acc, _ = orc.mesh_acc(nx, dx, x, want_abs=False)
    assert np.all(acc[0] < 0)
    with pytest.raises(AssertionError):
        np.testing.assert_allclose(acc[0], 1.0)


def test_acc_abs_pins():
    # acc_abs (p2mesh-kernel.cuh:289-298), norm3 of common.cuh:84-96:
    # Pythagorean quadruples give integers exactly, signs do not matter,
    # and the rest agrees with numpy.linalg.norm.
    quads = np.array([[1, 2, 2, 3], [2, 3, 6, 7], [1, 4, 8, 9], [2, 6, 9, 11],
                      [3, 4, 12, 13], [0, 0, 0, 0], [0, 0, -5, 5]], dtype=np.float64)
    v = np.ascontiguousarray(quads[:, :3].T)
    np.testing.assert_array_equal(orc.acc_abs(v), quads[:, 3])
    rng = np.random.default_rng(23)
    v = rng.normal(size=(3, 1000)) * 10.0 ** rng.integers(-8, 8, (1, 1000))
    out = orc.acc_abs(v)
    np.testing.assert_array_equal(orc.acc_abs(-v), out)
    np.testing.assert_allclose(out, np.linalg.norm(v, axis=0), rtol=4e-16)
