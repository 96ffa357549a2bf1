"""GPU parity of khrr_p2mesh_bin (p2mesh-kernel.cuh:11-137, the sequence of
nbody.cu:139-181) against the oracle: every output is compared bit for bit
(integer outputs and the index-order FP64 cell masses)."""
import numpy as np
import pytest

import oracle as orc
from helpers import galaxy_system, to_dev

pytestmark = pytest.mark.gpu


def _grid(nx, lo, hi):
    import khrr
    return khrr.cube_grid(nx, lo, hi)


def _bin_gpu(ctx, grid, pos, mass):
    import torch
    n = pos.shape[1]
    G = grid.size[0] * grid.size[1] * grid.size[2]
    B = (G + 511) // 512
    i32 = dict(device="cuda:0", dtype=torch.int32)
    out = dict(cell_id=torch.full((n,), -7, **i32), id_in_cell=torch.full((n,), -7, **i32),
               count=torch.full((G,), -7, **i32), offset=torch.full((G,), -7, **i32),
               start_id_ref=torch.full((G,), -7, **i32), block_incl=torch.full((B,), -7, **i32),
               cell_mass=torch.full((G,), -7.0, device="cuda:0", dtype=torch.float64))
    ctx.p2mesh_bin(grid, to_dev(pos, torch.float64), to_dev(mass, torch.float64), **out)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _bin_ref(grid, pos, mass):
    return orc.p2mesh_bin(list(grid.min), list(grid.cell), list(grid.size), pos, mass)


def _check_equal(got, ref):
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


@pytest.mark.parametrize("nx,lo,hi,n", [(64, -2.0, 2.0, 20000), (13, -1.5, 1.5, 3001),
                                        (200, -20.0, 20.0, 5000), (1, -1.0, 1.0, 100)])
def test_bit_exact_vs_oracle(ctx, nx, lo, hi, n):
    pos, _, mass, _ = galaxy_system(n // 3, n - n // 3, seed=nx)
    g = _grid(nx, lo, hi)
    _check_equal(_bin_gpu(ctx, g, pos, mass), _bin_ref(g, pos, mass))


def test_full_size_headline_grid(ctx):
    # 786432 bodies on 256^3 over [-2,2]^3 (the headline workload's binning)
    pos, _, mass, _ = galaxy_system(262144, 524288, seed=12345)
    g = _grid(256, -2.0, 2.0)
    got = _bin_gpu(ctx, g, pos, mass)
    ref = _bin_ref(g, pos, mass)
    _check_equal(got, ref)
    # size-independent properties
    assert got["count"].sum() == pos.shape[1]
    assert got["block_incl"][-1] == pos.shape[1]
    assert abs(got["cell_mass"].sum() - mass.sum()) < 1e-9 * mass.sum()


def test_edge_inputs(ctx):
    # NaN -> 0 then clamp, +-inf and huge values saturate then clamp,
    # everything outside lands on faces; many particles share one cell
    n = 5000
    rng = np.random.default_rng(3)
    pos = rng.uniform(-3, 3, (3, n))
    pos[:, :1000] = 0.01  # one dense cell
    pos[0, 1000] = np.nan
    pos[1, 1001] = np.inf
    pos[2, 1002] = -np.inf
    pos[:, 1003] = 1e300
    pos[:, 1004] = -1e300
    pos[0, 1005] = 2.0  # exactly the upper bound
    mass = rng.uniform(0.1, 1.0, n)
    g = _grid(16, -2.0, 2.0)
    _check_equal(_bin_gpu(ctx, g, pos, mass), _bin_ref(g, pos, mass))


def test_all_in_one_cell_and_empty(ctx):
    g = _grid(8, -1.0, 1.0)
    pos = np.full((3, 4000), 0.3)
    mass = np.linspace(1.0, 2.0, 4000)
    got = _bin_gpu(ctx, g, pos, mass)
    _check_equal(got, _bin_ref(g, pos, mass))
    np.testing.assert_array_equal(got["id_in_cell"], np.arange(4000))
    empty = _bin_gpu(ctx, g, np.zeros((3, 0)), np.zeros(0))
    assert np.all(empty["count"] == 0) and np.all(empty["offset"] == 0)
    assert np.all(empty["cell_mass"] == 0.0)


def test_dense_cells_across_mass_chunks(ctx):
    # three crowded cells of one 512-cell block, interleaved in index order,
    # with masses over many magnitudes: the block's masses are gathered in
    # chunks of 2048 whose boundaries cut through the cells, and each cell's
    # FP64 sum must still be the index-order one
    rng = np.random.default_rng(17)
    g = _grid(4, -1.0, 1.0)  # 64 cells, one block
    centres = np.array([[-0.75, -0.75, -0.75], [0.25, -0.25, 0.75], [0.75, 0.75, -0.25]])
    which = rng.choice(3, 7200, p=[0.35, 0.4, 0.25])
    pos = centres[which].T + rng.uniform(-0.2, 0.2, (3, which.size))
    mass = 10.0 ** rng.uniform(-12, 3, which.size)
    got = _bin_gpu(ctx, g, pos, mass)
    _check_equal(got, _bin_ref(g, pos, mass))
    assert np.sort(got["count"])[-3:].min() > 1500


def test_crowded_blocks_among_many(ctx):
    # 160^3 cells make 8000 blocks of 512, far more than stay resident at once;
    # every 256th block holds three crowded cells (2700 bodies, past the 2048
    # that fit in registers, so their masses go through the permutation
    # scratch) among sparsely occupied and empty blocks
    nx = 160
    g = _grid(nx, -2.0, 2.0)
    h = 4.0 / nx

    def centre(ijk):
        return -2.0 + (np.asarray(ijk, dtype=np.float64) + 0.5) * h

    probe = np.array([centre(v) for v in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 0, 0))]).T
    cid = _bin_ref(g, probe, np.ones(4))["cell_id"].astype(np.int64)
    strides = cid[:3] - cid[3]
    assert sorted(strides.tolist()) == [1, nx, nx * nx]
    rng = np.random.default_rng(23)
    parts = [rng.uniform(-2.0, 2.0, (3, 20000))]
    for blk in range(0, nx ** 3 // 512, 256):
        for c in blk * 512 + rng.choice(512, 3, replace=False):
            ijk = [(c // s) % nx for s in strides]
            parts.append(centre(ijk)[:, None] + rng.uniform(-0.3 * h, 0.3 * h, (3, 900)))
    pos = np.concatenate(parts, axis=1)
    pos = pos[:, rng.permutation(pos.shape[1])]
    mass = 10.0 ** rng.uniform(-6, 2, pos.shape[1])
    got = _bin_gpu(ctx, g, pos, mass)
    ref = _bin_ref(g, pos, mass)
    _check_equal(got, ref)
    assert (np.diff(got["block_incl"], prepend=0) > 2048).sum() >= 32


def test_fine_grid_outputs_at_odd_offset(ctx):
    # outputs at an odd element offset of their allocation on a fine grid
    import torch
    pos, _, mass, _ = galaxy_system(4000, 8000, seed=5)
    g = _grid(200, -2.0, 2.0)
    n, G = pos.shape[1], 200 ** 3
    cnt = torch.full((G + 1,), -7, device="cuda:0", dtype=torch.int32)
    cm = torch.full((G + 1,), -7.0, device="cuda:0", dtype=torch.float64)
    cell_id = torch.full((n,), -7, device="cuda:0", dtype=torch.int32)
    ctx.p2mesh_bin(g, to_dev(pos, torch.float64), to_dev(mass, torch.float64), cell_id=cell_id,
                   count=cnt[1:], cell_mass=cm[1:])
    torch.cuda.synchronize()
    ref = _bin_ref(g, pos, mass)
    np.testing.assert_array_equal(cnt[1:].cpu().numpy(), ref["count"])
    np.testing.assert_array_equal(cm[1:].cpu().numpy(), ref["cell_mass"])
    np.testing.assert_array_equal(cell_id.cpu().numpy(), ref["cell_id"])
    assert cnt[0].item() == -7 and cm[0].item() == -7.0


def test_reference_start_file_fixture(ctx):
    import os
    from conftest import GOLDEN
    gd = np.load(os.path.join(GOLDEN, "p2mesh_golden.npz"))
    pos = gd["pos"]
    mass = np.full(pos.shape[1], 1.0 / 262144)
    got200 = _bin_gpu(ctx, _grid(200, -20.0, 20.0), pos, mass)
    np.testing.assert_array_equal(got200["cell_id"], gd["cell_200"])
    got64 = _bin_gpu(ctx, _grid(64, -2.0, 2.0), pos, mass)
    np.testing.assert_array_equal(got64["cell_id"], gd["cell_64"])


def test_rejects_bad_grid(ctx):
    import khrr
    g = _grid(4, -1.0, 1.0)
    g.size[0] = 0
    with pytest.raises(khrr.KhrrError, match="invalid grid"):
        _bin_gpu(ctx, g, np.zeros((3, 1)), np.ones(1))


@pytest.mark.parametrize("nx,n", [(24, 20000), (128, 200000), (5, 1)])
def test_cell_transfers_bit_exact(ctx, nx, n):
    # computeCellPhiUnsorted, apply_particle_to_cell_weighted and
    # apply_cell_to_particles_avg (p2mesh-kernel.cuh:152-167,256-287) on the
    # outputs of a binning: index-order FP64 sums, bit for bit.
    import torch
    pos, _, mass, _ = galaxy_system(max(1, n // 3), n - n // 3, seed=nx) if n > 1 else (
        np.zeros((3, 1)), None, np.array([0.5]), None)
    n = pos.shape[1]
    g = _grid(nx, -2.0, 2.0)
    G = nx ** 3
    got = _bin_gpu(ctx, g, pos, mass)
    ref = _bin_ref(g, pos, mass)
    _check_equal(got, ref)
    d = {k: to_dev(v, torch.int32) for k, v in got.items() if k in (
        "cell_id", "id_in_cell", "count", "offset")}
    rng = np.random.default_rng(nx)
    f64 = dict(device="cuda:0", dtype=torch.float64)
    # cell potential of the bodies, fresh and accumulated
    phi = (rng.normal(size=n) * 3.0).astype(np.float32)
    cphi = torch.full((G,), 5.0, **f64)
    ctx.p2mesh_cell_phi(d["cell_id"], d["id_in_cell"], d["count"], d["offset"],
                        to_dev(phi, torch.float32), cphi)
    ref_cphi = orc.p2mesh_cell_phi(ref["cell_id"], ref["count"], phi.astype(np.float64))
    np.testing.assert_array_equal(cphi.cpu().numpy(), ref_cphi)
    start = rng.normal(size=G)
    acc = to_dev(start, torch.float64)
    ctx.p2mesh_cell_phi(d["cell_id"], d["id_in_cell"], d["count"], d["offset"],
                        to_dev(phi, torch.float32), acc, accumulate=True)
    np.testing.assert_array_equal(acc.cpu().numpy(), orc.p2mesh_cell_phi(
        ref["cell_id"], ref["count"], phi.astype(np.float64), cell_phi=start, accumulate=True))
    # mass-weighted cell mean of a body value
    value = rng.normal(size=n) * 7.0
    cm = to_dev(got["cell_mass"], torch.float64)
    cval = torch.full((G,), -3.0, **f64)
    ctx.p2mesh_cell_weighted(d["cell_id"], d["id_in_cell"], d["count"], d["offset"], cm,
                             to_dev(mass, torch.float64), to_dev(value, torch.float64), cval)
    ref_w = orc.p2mesh_cell_weighted(ref["cell_id"], ref["cell_mass"], mass, value)
    np.testing.assert_array_equal(cval.cpu().numpy(), ref_w)
    # cell value shared out over the bodies
    out = torch.full((n,), -3.0, **f64)
    ctx.p2mesh_cell_avg(d["cell_id"], d["count"], cm, out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), orc.p2mesh_cell_avg(
        ref["cell_id"], ref["count"], ref["cell_mass"]))


def test_cell_avg_empty_and_bad_cells(ctx):
    import torch
    G = 10
    cid = np.array([0, 3, 3, -1, 10, 7], np.int32)
    count = np.array([1, 0, 0, 2, 0, 0, 0, 0, 0, 0], np.int32)
    vals = np.arange(G, dtype=np.float64) + 0.5
    out = torch.zeros(cid.size, device="cuda:0", dtype=torch.float64)
    ctx.p2mesh_cell_avg(to_dev(cid, torch.int32), to_dev(count, torch.int32),
                        to_dev(vals, torch.float64), out)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert o[0] == 0.5 and o[1] == 3.5 / 2 and o[2] == 3.5 / 2
    assert np.isnan(o[3]) and np.isnan(o[4])
    assert o[5] == -1.0
