"""GPU mesh force (khrr_mesh_acc, khrr_mesh_to_particles) against the FP64
oracle: the FP32 field is widened, differenced and scaled in FP64 and
rounded once, so the acceleration field matches the oracle rounded to FP32
bit for bit; the whole wave-solver chain (bin -> source -> iterate -> mesh
acceleration -> particles) matches the FP64 chain within the stencil's
1e-5 tolerance."""
import numpy as np
import pytest

import oracle as orc
from helpers import to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx", [2, 37, 64, 70])
def test_field_bit_exact(ctx, nx):
    import torch
    G = nx ** 3
    dx = 4.0 / nx
    phi = np.random.default_rng(nx).normal(size=G).astype(np.float32) * 25.0
    acc = torch.zeros((3, G), device="cuda:0")
    ab = torch.zeros(G, device="cuda:0")
    ctx.mesh_acc(nx, dx, to_dev(phi, torch.float32), acc, ab)
    torch.cuda.synchronize()
    ref, ref_ab = orc.mesh_acc(nx, dx, phi.astype(np.float64))
    np.testing.assert_array_equal(acc.cpu().numpy(), ref.astype(np.float32))
    np.testing.assert_allclose(ab.cpu().numpy(), ref_ab.astype(np.float32), rtol=1.2e-7)


def test_gather_and_bad_ids(ctx):
    import torch
    rng = np.random.default_rng(9)
    G = 1000
    cell_acc = rng.normal(size=(3, G)).astype(np.float32)
    cell_phi = rng.normal(size=G).astype(np.float32)
    cid = rng.integers(0, G, 4096).astype(np.int32)
    cid[7] = -1
    cid[8] = G
    acc = torch.zeros((3, cid.size), device="cuda:0")
    phi = torch.zeros(cid.size, device="cuda:0")
    ctx.mesh_to_particles(to_dev(cid, torch.int32), to_dev(cell_acc, torch.float32),
                          to_dev(cell_phi, torch.float32), acc, phi)
    torch.cuda.synchronize()
    a, p = acc.cpu().numpy(), phi.cpu().numpy()
    ok = (cid >= 0) & (cid < G)
    np.testing.assert_array_equal(a[:, ok], cell_acc[:, cid[ok]])
    np.testing.assert_array_equal(p[ok], cell_phi[cid[ok]])
    assert np.isnan(a[:, ~ok]).all() and np.isnan(p[~ok]).all()


def test_rejects_bad_arguments(ctx):
    import khrr
    import torch
    f = torch.zeros(8, device="cuda:0")
    with pytest.raises(khrr.KhrrError):
        ctx.mesh_acc(1, 1.0, torch.zeros(1, device="cuda:0"), torch.zeros((3, 1), device="cuda:0"))
    with pytest.raises(khrr.KhrrError):
        ctx.mesh_acc(2, 0.0, f, torch.zeros((3, 8), device="cuda:0"))


def test_wave_solver_chain(ctx):
    # calc_acceleration_by_grid (nbody.cu:183-285): bin the bodies, iterate
    # the damped wave equation, take -grad(phi) and read it at the bodies
    import khrr
    import torch
    from helpers import galaxy_system
    pos, _, mass, _ = galaxy_system(3000, 6000, seed=12, eps_s=0.004, eps_d=0.004)
    n = pos.shape[1]
    nx, iters = 40, 120
    dx = 4.0 / nx
    grid = khrr.cube_grid(nx, -2.0, 2.0)
    G = nx ** 3
    i32 = dict(device="cuda:0", dtype=torch.int32)
    cell_id = torch.zeros(n, **i32)
    count = torch.zeros(G, **i32)
    cm = torch.zeros(G, device="cuda:0", dtype=torch.float64)
    ctx.p2mesh_bin(grid, to_dev(pos, torch.float64), to_dev(mass, torch.float64),
                   cell_id=cell_id, count=count, cell_mass=cm)
    w = khrr.reference_wave(nx, dx)
    src = torch.zeros(G, device="cuda:0")
    ctx.wave_source(w, cm, src)
    b = [torch.zeros(G, device="cuda:0") for _ in range(3)]
    _, cur, _ = ctx.wave_iterate(w, b[0], b[1], b[2], src, iters)
    cacc = torch.zeros((3, G), device="cuda:0")
    ctx.mesh_acc(nx, dx, cur, cacc)
    acc = torch.zeros((3, n), device="cuda:0")
    phi = torch.zeros(n, device="cuda:0")
    ctx.mesh_to_particles(cell_id, cacc, cur, acc, phi)
    torch.cuda.synchronize()
    wd = orc.reference_wave(nx, dx)
    ref_phi, _ = orc.wave_iterate(wd, np.zeros(G), np.zeros(G), cm.cpu().numpy(), iters)
    ref_cacc, _ = orc.mesh_acc(nx, dx, ref_phi, want_abs=False)
    ref_acc, ref_p = orc.mesh_to_particles(cell_id.cpu().numpy(), ref_cacc, ref_phi)
    a = acc.cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(ref_acc))
    assert scale > 0
    assert np.max(np.abs(a - ref_acc)) / scale <= 1e-4
    assert np.max(np.abs(phi.cpu().numpy() - ref_p)) / np.max(np.abs(ref_p)) <= 1e-5


@pytest.mark.parametrize("nx", [2, 37, 70])
def test_field_f64_bit_exact(ctx, nx):
    # the FP64 form (the debug dumps' differencing of the bodies' FP64 cell
    # potential): the same expression as the oracle in FP64; |a| within one
    # FP64 ulp (device sqrt)
    import torch
    G = nx ** 3
    dx = 4.0 / nx
    phi = np.random.default_rng(nx + 1).normal(size=G) * 25.0
    acc = torch.zeros((3, G), device="cuda:0", dtype=torch.float64)
    ab = torch.zeros(G, device="cuda:0", dtype=torch.float64)
    ctx.mesh_acc_f64(nx, dx, to_dev(phi, torch.float64), acc, ab)
    torch.cuda.synchronize()
    ref, ref_ab = orc.mesh_acc(nx, dx, phi)
    np.testing.assert_array_equal(acc.cpu().numpy(), ref)
    np.testing.assert_array_max_ulp(ab.cpu().numpy(), ref_ab, maxulp=1)


def test_acc_abs(ctx):
    # acc_abs (p2mesh-kernel.cuh:289-298) of FP32 planes in FP64: Pythagorean
    # quadruples exact, the rest within one FP64 ulp of the oracle
    import torch
    rng = np.random.default_rng(31)
    v = (rng.normal(size=(3, 5000)) * 10.0 ** rng.integers(-6, 6, (1, 5000))).astype(np.float32)
    v[:, :3] = [[1, 2, 3], [2, 3, 4], [2, 6, 12]]
    out = torch.zeros(v.shape[1], device="cuda:0", dtype=torch.float64)
    ctx.acc_abs(to_dev(v, torch.float32), out)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert list(o[:3]) == [3.0, 7.0, 13.0]
    np.testing.assert_array_max_ulp(o, orc.acc_abs(v.astype(np.float64)), maxulp=1)


def test_debug_transfers_reject_bad_arguments(ctx):
    # the entry points behind the debug dumps fail with KHRR_EINVAL, not a
    # launch, on bad sizes and missing buffers
    import khrr
    import torch
    lib = khrr.lib()
    f64 = torch.zeros(8, device="cuda:0", dtype=torch.float64)
    i32 = torch.zeros(8, device="cuda:0", dtype=torch.int32)
    p = lambda t: t.data_ptr()
    EINVAL = -1  # KHRR_EINVAL
    assert lib.khrr_acc_abs(ctx._h, -1, None, None) == EINVAL
    assert lib.khrr_acc_abs(ctx._h, 4, None, p(f64)) == EINVAL
    assert lib.khrr_acc_abs(ctx._h, 0, None, None) == 0
    assert lib.khrr_mesh_acc_f64(ctx._h, 1, 1.0, p(f64), p(f64), None) == EINVAL
    assert lib.khrr_mesh_acc_f64(ctx._h, 2, -1.0, p(f64), p(f64), None) == EINVAL
    assert lib.khrr_mesh_acc_f64(ctx._h, 2, 1.0, None, p(f64), None) == EINVAL
    assert lib.khrr_p2mesh_cell_avg(ctx._h, 4, 0, p(i32), p(i32), p(f64), p(f64)) == EINVAL
    assert lib.khrr_p2mesh_cell_avg(ctx._h, 4, 8, None, p(i32), p(f64), p(f64)) == EINVAL
    assert lib.khrr_p2mesh_cell_weighted(ctx._h, -1, 8, p(i32), p(i32), p(i32), p(i32), p(f64),
                                         p(f64), p(f64), p(f64)) == EINVAL
    assert lib.khrr_p2mesh_cell_weighted(ctx._h, 4, 8, p(i32), p(i32), p(i32), p(i32), p(f64),
                                         None, p(f64), p(f64)) == EINVAL
    assert "null" in lib.khrr_last_error(ctx._h).decode()
