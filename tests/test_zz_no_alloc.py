"""khrr_p2mesh_bin_noalloc, the binning entry point that allocates nothing:
through the C-ABI its outputs must equal khrr_p2mesh_bin's bit for bit on
two calls in a row (the second smaller, so the fixed scratch still holds
the first call's contents), and sizes beyond its fixed scratch are refused
with a clear error before any launch."""
import numpy as np
import pytest

from helpers import galaxy_system

pytestmark = pytest.mark.gpu


def _outputs(dev, n, G, fill):
    import torch
    i32 = dict(device=dev, dtype=torch.int32)
    B = (G + 511) // 512
    return dict(cell_id=torch.full((n,), fill, **i32), id_in_cell=torch.full((n,), fill, **i32),
                count=torch.full((G,), fill, **i32), offset=torch.full((G,), fill, **i32),
                start_id_ref=torch.full((G,), fill, **i32),
                block_incl=torch.full((B,), fill, **i32),
                cell_mass=torch.full((G,), float(fill), device=dev, dtype=torch.float64))


def _bin_both(ctx, dev, grid, pos, mass):
    import torch
    n = pos.shape[1]
    G = grid.size[0] * grid.size[1] * grid.size[2]
    P = torch.tensor(pos, device=dev, dtype=torch.float64)
    M = torch.tensor(mass, device=dev, dtype=torch.float64)
    ref = _outputs(dev, n, G, -7)
    ctx.p2mesh_bin(grid, P, M, **ref)
    got = _outputs(dev, n, G, -9)  # stale values would show as -9
    print(f"noalloc binning: n={n} cells={G}", flush=True)
    ctx.p2mesh_bin_noalloc(grid, P, M, **got)
    torch.cuda.synchronize()
    for k in ref:
        np.testing.assert_array_equal(got[k].cpu().numpy(), ref[k].cpu().numpy(), err_msg=k)
    return {k: v.cpu().numpy() for k, v in got.items()}


def test_noalloc_equals_allocating_form(ctx, dev):
    import khrr
    # first call: the headline galaxy (786432 bodies) on 256^3 cells, the
    # largest grid the fixed scratch takes
    pos, _, mass, _ = galaxy_system(262144, 524288, seed=12345)
    out = _bin_both(ctx, dev, khrr.cube_grid(256, -2.0, 2.0), pos, mass)
    assert out["count"].sum() == pos.shape[1]
    assert abs(out["cell_mass"].sum() - mass.sum()) <= 1e-12 * mass.sum()
    # second call, smaller: a different galaxy on 64^3 cells
    pos2, _, mass2, _ = galaxy_system(3000, 9000, seed=77)
    out2 = _bin_both(ctx, dev, khrr.cube_grid(64, -2.0, 2.0), pos2, mass2)
    assert out2["count"].sum() == pos2.shape[1]
    # and a third, ragged one: 1 body on a single cell
    g1 = khrr.cube_grid(1, -2.0, 2.0)
    out3 = _bin_both(ctx, dev, g1, pos2[:, :1].copy(), mass2[:1].copy())
    assert out3["count"].tolist() == [1] and out3["id_in_cell"].tolist() == [0]


def test_noalloc_refuses_sizes_beyond_its_scratch(ctx, dev):
    import torch
    import khrr
    # 2^20 + 1 bodies: refused before anything runs
    n = (1 << 20) + 1
    P = torch.zeros((3, n), device=dev, dtype=torch.float64)
    M = torch.zeros(n, device=dev, dtype=torch.float64)
    g = khrr.cube_grid(16, -2.0, 2.0)
    cnt = torch.zeros(16 ** 3, device=dev, dtype=torch.int32)
    with pytest.raises(khrr.KhrrError, match="exceed its fixed scratch"):
        ctx.p2mesh_bin_noalloc(g, P, M, count=cnt)
    # 257^3 cells: refused too
    g2 = khrr.cube_grid(257, -2.0, 2.0)
    cnt2 = torch.zeros(257 ** 3, device=dev, dtype=torch.int32)
    with pytest.raises(khrr.KhrrError, match="exceed its fixed scratch"):
        ctx.p2mesh_bin_noalloc(g2, P[:, :10].contiguous(), M[:10].contiguous(), count=cnt2)
    # the allocating form still takes what the fixed scratch does not
    cnt3 = torch.zeros(16 ** 3, device=dev, dtype=torch.int32)
    ctx.p2mesh_bin(g, P, M, count=cnt3)
    torch.cuda.synchronize()
    assert int(cnt3.sum()) == n
