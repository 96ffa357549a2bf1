"""Pins the oracle's binning (p2mesh-kernel.cuh:11-137) with an independent
numpy restatement, brute-force recounts and the golden cell indices."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN


def _numpy_cells(pos, lo, d, s):
    def axis(v, size):
        q = (v - lo) / d
        with np.errstate(invalid="ignore"):
            k = np.where(np.isnan(q), 0, np.clip(q, -2.0 ** 31, 2.0 ** 31 - 1))
        k = np.trunc(k).astype(np.int64)
        return np.clip(k, 0, size - 1)
    ix, iy, iz = axis(pos[0], s), axis(pos[1], s), axis(pos[2], s)
    return (iz * s * s + iy * s + ix).astype(np.int32)


def _random(n, seed, spread=2.5):
    rng = np.random.default_rng(seed)
    return rng.normal(scale=spread / 2, size=(3, n)), rng.uniform(0.1, 1.0, n)


@pytest.mark.parametrize("s,lo,hi", [(16, -2.0, 2.0), (13, -1.0, 1.0), (64, -2.0, 2.0)])
def test_against_numpy_recount(s, lo, hi):
    pos, m = _random(20000, s)
    d = (hi - lo) / s
    out = orc.p2mesh_bin([lo] * 3, [d] * 3, [s] * 3, pos, m)
    cells = _numpy_cells(pos, lo, d, s)
    np.testing.assert_array_equal(out["cell_id"], cells)
    G = s ** 3
    count = np.bincount(cells, minlength=G)
    np.testing.assert_array_equal(out["count"], count)
    # exclusive per-cell scan
    np.testing.assert_array_equal(out["offset"], np.concatenate([[0], np.cumsum(count)[:-1]]))
    # block-granular offsets of the reference (computePrefixSums +
    # adjustGlobalPrefixSums): sum of counts of all 512-blocks before
    B = (G + 511) // 512
    blk = np.array([count[b * 512:(b + 1) * 512].sum() for b in range(B)])
    incl = np.cumsum(blk)
    np.testing.assert_array_equal(out["block_incl"], incl)
    excl = np.concatenate([[0], incl[:-1]])
    np.testing.assert_array_equal(out["start_id_ref"], excl[np.arange(G) // 512])
    # id_in_cell is the index-order rank: a permutation of 0..count-1
    for c in np.unique(cells)[:200]:
        members = np.nonzero(cells == c)[0]
        np.testing.assert_array_equal(out["id_in_cell"][members], np.arange(len(members)))
    # cell mass: index-order FP64 sum, and mass conservation
    for c in np.unique(cells)[:200]:
        acc = 0.0
        for i in np.nonzero(cells == c)[0]:
            acc += m[i]
        assert out["cell_mass"][c] == acc
    assert abs(out["cell_mass"].sum() - m.sum()) < 1e-12 * m.sum()


def test_saturation_and_nan():
    assert orc.sat_trunc(float("nan")) == 0
    assert orc.sat_trunc(1e300) == 2 ** 31 - 1
    assert orc.sat_trunc(-1e300) == -2 ** 31
    assert orc.sat_trunc(2147483647.9) == 2147483647
    assert orc.sat_trunc(-2147483648.9) == -2147483648
    assert orc.sat_trunc(-0.9) == 0 and orc.sat_trunc(3.99) == 3
    pos = np.array([[np.nan, np.inf, -np.inf, 1e300], [0, 0, 0, 0], [0, 0, 0, 0.0]])
    out = orc.p2mesh_bin([-1] * 3, [0.5] * 3, [4] * 3, pos, np.ones(4))
    ix = out["cell_id"] % 4
    np.testing.assert_array_equal(ix, [0, 3, 0, 3])


def test_golden_cells_of_reference_stars():
    g = np.load(os.path.join(GOLDEN, "p2mesh_golden.npz"))
    pos = g["pos"]
    m = np.ones(pos.shape[1])
    o200 = orc.p2mesh_bin([-20.0] * 3, [0.2] * 3, [200] * 3, pos, m)
    np.testing.assert_array_equal(o200["cell_id"], g["cell_200"])
    o64 = orc.p2mesh_bin([-2.0] * 3, [4.0 / 64] * 3, [64] * 3, pos, m)
    np.testing.assert_array_equal(o64["cell_id"], g["cell_64"])


def test_invalid_grids_rejected():
    pos, m = _random(10, 0)
    with pytest.raises(ValueError):
        orc.p2mesh_bin([0] * 3, [1] * 3, [0, 4, 4], pos, m)
    with pytest.raises(ValueError):
        orc.p2mesh_bin([0] * 3, [1] * 3, [1, 4, 4], pos, m)  # index escapes the grid


def test_empty_input():
    out = orc.p2mesh_bin([-1] * 3, [0.5] * 3, [4] * 3, np.zeros((3, 0)), np.zeros(0))
    assert out["count"].sum() == 0 and np.all(out["offset"] == 0)


def test_cell_phi_pins():
    # computeCellPhiUnsorted (p2mesh-kernel.cuh:152-167): every particle adds
    # phi_i / count[cell] to its cell, so a cell holds the mean potential of
    # its particles.  Pinned by what that fixes: a one-particle cell gets its
    # phi exactly, a cell of equal phi and a power-of-two count gets that
    # value exactly, count-weighted cell values sum to sum(phi), empty cells
    # stay 0 (or keep their value when accumulating), and accumulating adds
    rng = np.random.default_rng(12)
    G = 40
    cell_id = np.concatenate([
        np.full(4, 3), np.full(8, 7), np.array([11]), np.array([19]),
        rng.integers(20, 36, 500)]).astype(np.int32)
    phi = rng.normal(size=cell_id.shape[0])
    phi[:4] = -1.75   # cell 3: four equal values
    phi[4:12] = 0.3   # cell 7: eight equal values
    count = np.bincount(cell_id, minlength=G).astype(np.int32)
    out = orc.p2mesh_cell_phi(cell_id, count, phi)
    assert out[3] == -1.75
    assert out[7] == 0.3 * 8 / 8 or abs(out[7] - 0.3) <= 4e-16
    assert out[11] == phi[12] and out[19] == phi[13]
    empty = count == 0
    assert np.all(out[empty] == 0.0)
    np.testing.assert_allclose(np.sum(out * count), np.sum(phi), rtol=1e-12)
    for c in range(20, 36):  # mean per cell
        np.testing.assert_allclose(out[c], np.mean(phi[cell_id == c]), rtol=1e-13, atol=1e-15)
    start = rng.normal(size=G)
    acc = orc.p2mesh_cell_phi(cell_id, count, phi, cell_phi=start, accumulate=True)
    np.testing.assert_allclose(acc, start + out, rtol=1e-14, atol=1e-15)
    assert np.all(acc[empty] == start[empty])


def test_cell_weighted_pins():
    # apply_particle_to_cell_weighted (p2mesh-kernel.cuh:273-287): every body
    # adds value_i * m_i / M_cell, so a cell holds the mass-weighted mean of
    # its bodies' values.  Pinned by what that fixes: one-body cells get the
    # value exactly, four bodies of mass 1/4 sum exactly, a constant value is
    # returned, sum_c M_c v_c = sum_i m_i v_i, empty cells stay 0, and the
    # mean equals numpy's weighted average.  Using counts in place of masses,
    # or another cell's mass, fails the conservation and average checks.
    rng = np.random.default_rng(21)
    G = 50
    cell_id = np.concatenate([
        np.full(4, 2), np.array([5]), np.array([9]),
        rng.integers(10, 40, 700)]).astype(np.int32)
    n = cell_id.shape[0]
    mass = rng.uniform(0.1, 3.0, n)
    mass[:4] = 0.25
    value = rng.normal(size=n)
    value[:4] = [1.0, 2.0, 3.0, 4.0]
    cell_mass = np.zeros(G)
    np.add.at(cell_mass, cell_id, mass)
    out = orc.p2mesh_cell_weighted(cell_id, cell_mass, mass, value)
    assert out[2] == 2.5
    assert out[5] == value[4] and out[9] == value[5]
    empty = np.bincount(cell_id, minlength=G) == 0
    assert np.all(out[empty] == 0.0)
    np.testing.assert_allclose(np.sum(cell_mass * out), np.sum(mass * value), rtol=1e-12)
    for c in range(10, 40):
        sel = cell_id == c
        if sel.any():
            np.testing.assert_allclose(out[c], np.average(value[sel], weights=mass[sel]),
                                       rtol=1e-12, atol=1e-14)
    const = orc.p2mesh_cell_weighted(cell_id, cell_mass, mass, np.full(n, -0.7))
    np.testing.assert_allclose(const[~empty], -0.7, rtol=1e-14)
    # linear in the value
    v2 = rng.normal(size=n)
    both = orc.p2mesh_cell_weighted(cell_id, cell_mass, mass, 2.0 * value + v2)
    np.testing.assert_allclose(both, 2.0 * out + orc.p2mesh_cell_weighted(
        cell_id, cell_mass, mass, v2), rtol=1e-12, atol=1e-13)


def test_cell_avg_pins():
    # apply_cell_to_particles_avg (p2mesh-kernel.cuh:256-271): the cell's
    # value divided by its count at each body, -1 where the count is 0.
    # Equal masses of 1/4 come back exactly from the binning's cell mass.
    rng = np.random.default_rng(22)
    G = 30
    cell_id = rng.integers(0, 20, 400).astype(np.int32)
    count = np.bincount(cell_id, minlength=G).astype(np.int32)
    mass = np.full(cell_id.shape[0], 0.25)
    cell_mass = np.zeros(G)
    np.add.at(cell_mass, cell_id, mass)
    out = orc.p2mesh_cell_avg(cell_id, count, cell_mass)
    assert np.all(out == 0.25)
    vals = rng.normal(size=G)
    out = orc.p2mesh_cell_avg(cell_id, count, vals)
    np.testing.assert_array_equal(out, vals[cell_id] / count[cell_id])
    bad = count.copy()
    bad[cell_id[0]] = 0
    out = orc.p2mesh_cell_avg(cell_id, bad, vals)
    assert np.all(out[cell_id == cell_id[0]] == -1.0)
