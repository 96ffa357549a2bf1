"""The serving form (khrr_server_*): continuous batching with admission by
device-memory footprint.  Each result must be bit-identical whether the
request ran alone or among others; energies follow the FP64 oracle loop
(nbody.cu:1468-1594) within FP32-force tolerance (1e-5 relative)."""
import ctypes

import numpy as np
import pytest

import ics
import oracle as orc

pytestmark = pytest.mark.gpu


def _plummer_req(n, seed):
    pos, vel = ics.plummer(n, 1.0, 0.3333, 1.0, seed=seed)
    eps = 0.004 * (n / 16384.0) ** (-1.0 / 3.0)
    return pos, vel, np.full(n, 1.0 / n), np.full(n, eps)


def _galaxy_req(n_s, n_d, seed):
    ps, vs, pd, vd = ics.galaxy(n_s, n_d, seed=seed)
    pos, vel, mass, _ = ics.system_arrays(ps, vs, pd, vd, 1.0, 6.699356, 0.004, 0.004)
    return pos, vel, mass, np.full(pos.shape[1], 0.004)


def _grid():
    import khrr
    return khrr.cube_grid(16, -2.0, 2.0)


def _run_all(srv, reqs, arrivals):
    """Submits request k before round arrivals[k]; steps until all finish."""
    ids = {}
    done = {}
    rnd = 0
    while len(done) < len(reqs):
        for k, r in enumerate(reqs):
            if k not in ids and arrivals[k] <= rnd:
                ids[k] = srv.submit(*r["arrays"], 0.002, r["steps"], _grid())
        for fid in srv.step():
            k = next(k for k, v in ids.items() if v == fid)
            done[k] = srv.result(fid, 16 ** 3)
        rnd += 1
        assert rnd < 10000
    return done


@pytest.mark.parametrize("kind", ["galaxy", "large"])
def test_alone_equals_in_the_mix(ctx, kind):
    import khrr
    # "large" runs on the pair-symmetric path (n >= 16384), "galaxy" on the
    # one-sided one; the others span both
    if kind == "galaxy":
        target = dict(arrays=_galaxy_req(1024, 2048, 5), steps=12)
    else:
        pos, vel, mass, eps = _plummer_req(36000, 6)
        target = dict(arrays=(pos, vel, mass, np.full(36000, 0.004)), steps=6)
    others = [dict(arrays=_plummer_req(n, 40 + n), steps=s)
              for n, s in ((4096, 20), (1024, 5), (2048, 9), (40960, 4), (8192, 15))]
    alone = khrr.Server(ctx, 1 << 30, 8)
    r_alone = _run_all(alone, [target], [0])[0]
    alone.close()
    mix = khrr.Server(ctx, 1 << 30, 8)
    reqs = others[:2] + [target] + others[2:]
    r_mix = _run_all(mix, reqs, [0, 0, 3, 4, 5, 6])[2]
    mix.close()
    for a, b in zip(r_alone[:2], r_mix[:2]):
        assert a == b  # diagnostics dicts, exact
    np.testing.assert_array_equal(r_alone[2], r_mix[2])
    total = 1.0 + 6.699356 if kind == "galaxy" else 1.0
    assert abs(r_mix[2].sum() - total) < 1e-9


def test_energy_follows_oracle(ctx):
    import khrr
    pos, vel, mass, eps = _plummer_req(2048, 77)
    srv = khrr.Server(ctx, 1 << 28, 4)
    res = _run_all(srv, [dict(arrays=(pos, vel, mass, eps), steps=25)], [0])[0]
    srv.close()
    h = (0.5 * eps * eps).astype(np.float32).astype(np.float64)
    m32 = mass.astype(np.float32).astype(np.float64)
    _, _, d = orc.run(pos, vel, m32, h, 0.002, 25)
    E = lambda dd: 0.5 * (dd["Ek"] + dd["Ep"])  # noqa: E731
    # the server uses FP64 masses in the sums and FP32 masses in gravity
    assert abs(E(res[0]) / (0.5 * (d[0, 0] + d[0, 1])) - 1) < 1e-5
    assert abs(E(res[1]) / (0.5 * (d[-1, 0] + d[-1, 1])) - 1) < 1e-5


def test_admission_respects_budget(ctx):
    import khrr
    # an 8192-body system holds 152 B per body plus 8 j-pieces of partial
    # sums (128 B per body): 2293760 B, so a 5 MB budget holds two of them
    srv = khrr.Server(ctx, 5_000_000, 16)
    try:
        reqs = [dict(arrays=_plummer_req(8192, 90 + k), steps=4) for k in range(4)]
        ids = [srv.submit(*r["arrays"], 0.002, r["steps"], _grid()) for r in reqs]
        done = srv.step()
        st = srv.stats()
        assert st["active"] == 2 and st["pending"] == 2 and not done
        assert st["bytes_in_use"] == 2 * 2293760 <= st["bytes_budget"]
        finished = []
        for _ in range(100):
            finished += srv.step()
            if len(finished) == 4:
                break
        assert sorted(finished) == sorted(ids)
        for rid in ids:
            d0, d1, cm = srv.result(rid, 16 ** 3)
            assert d0["M"] == pytest.approx(1.0) and cm.sum() == pytest.approx(1.0)
    finally:
        srv.close()


def test_server_outlived_by_nothing(ctx):
    # closing a context closes its servers first; a server closed that way
    # can be closed again harmlessly
    import khrr
    c = khrr.Context(0)
    srv = khrr.Server(c, 1 << 24, 2)
    srv.submit(*_plummer_req(1024, 1), 0.002, 2, _grid())
    c.close()
    srv.close()
    assert srv._h is None


def test_errors(ctx):
    import khrr
    srv = khrr.Server(ctx, 1_000_000, 2)
    big = _plummer_req(8192, 3)
    with pytest.raises(khrr.KhrrError, match="never fit"):
        srv.submit(*big, 0.002, 3, _grid())
    small = _plummer_req(1024, 4)
    rid = srv.submit(*small, 0.002, 3, _grid())
    with pytest.raises(khrr.KhrrError, match="not finished"):
        srv.result(rid, 16 ** 3)
    with pytest.raises(khrr.KhrrError, match="unknown request"):
        srv.result(12345, 16 ** 3)
    # a grid khrr_p2mesh_bin refuses (16 x 32 x 16: the reference's cell
    # index, z clamped by size[1] and y strided by size[2], can pass the last
    # cell) is refused at submit, not when the simulation finishes
    bad = khrr.Grid((ctypes.c_double * 3)(-2.0, -2.0, -2.0), (ctypes.c_double * 3)(0.25, 0.125, 0.25),
                    (ctypes.c_int * 3)(16, 32, 16))
    with pytest.raises(khrr.KhrrError, match="can exceed the cell count"):
        srv.submit(*small, 0.002, 3, bad)
    assert srv.stats()["pending"] == 1
    # the queued request still runs to the end and is collected
    for _ in range(20):
        if srv.step():
            break
    d0, d1, cm = srv.result(rid, 16 ** 3)
    assert cm.sum() == pytest.approx(1.0)
    srv.close()


def test_round_schedule_changes_timing_not_results(ctx):
    import khrr
    reqs = [dict(arrays=_plummer_req(n, 500 + n), steps=s)
            for n, s in ((8192, 12), (1024, 30), (36864, 5), (2048, 8))]
    runs = []
    for pairs in (0.0, 1.0):  # lock-step; then least remaining work, one system per round
        srv = khrr.Server(ctx, 1 << 30, 8, round_pairs=pairs)
        ids = [srv.submit(*r["arrays"], 0.002, r["steps"], _grid()) for r in reqs]
        order, res = [], {}
        while len(res) < len(reqs):
            for fid in srv.step():
                order.append(ids.index(fid))
                res[ids.index(fid)] = srv.result(fid, 16 ** 3)
        st = srv.stats()
        srv.close()
        runs.append((order, res, st["rounds"]))
    (o0, r0, n0), (o1, r1, n1) = runs
    for k in range(len(reqs)):
        assert r0[k][0] == r1[k][0] and r0[k][1] == r1[k][1]
        np.testing.assert_array_equal(r0[k][2], r1[k][2])
    # one system per round: every step is its own round, least work first
    assert n1 == sum(r["steps"] for r in reqs) and n0 == max(r["steps"] for r in reqs)
    work = [r["arrays"][0].shape[1] ** 2 * r["steps"] for r in reqs]
    assert o1 == sorted(range(len(reqs)), key=lambda k: work[k])


def test_round_pairs_rejects_negative(ctx):
    import khrr
    srv = khrr.Server(ctx, 1 << 26, 2)
    with pytest.raises(khrr.KhrrError):
        srv.set_round_pairs(-1.0)
    srv.close()


def test_c4_sized_requests_alone_equal_mix(ctx):
    """The serving workload's request shapes at full size: a disk + halo
    galaxy of 65536 + 131072 bodies and a 131072-body Plummer sphere among
    small spheres, each returning its FP64 mass on a 128^3 grid over
    [-2, 2]^3.  Each large result equals the same request run alone, bit for
    bit; every grid holds its request's mass; a small request's energy at
    t = 0 and at the end follows the FP64 oracle loop."""
    import khrr
    grid = khrr.cube_grid(128, -2.0, 2.0)
    G = 128 ** 3

    def run(srv, reqs, arrivals):
        ids, done, rnd = {}, {}, 0
        while len(done) < len(reqs):
            for k, r in enumerate(reqs):
                if k not in ids and arrivals[k] <= rnd:
                    ids[k] = srv.submit(*r["arrays"], 0.002, r["steps"], grid)
            for fid in srv.step():
                k = next(k for k, v in ids.items() if v == fid)
                done[k] = srv.result(fid, G)
            rnd += 1
            assert rnd < 1000
        return done

    galaxy = dict(arrays=_galaxy_req(65536, 131072, 200001), steps=8)
    sphere = dict(arrays=_plummer_req(131072, 200002), steps=6)
    small = dict(arrays=_plummer_req(2048, 200003), steps=10)
    others = [dict(arrays=_plummer_req(n, 200010 + n), steps=s)
              for n, s in ((1024, 12), (5120, 7), (20480, 5))]
    budget = 8 << 30
    alone = {}
    for name, r in (("galaxy", galaxy), ("sphere", sphere)):
        srv = khrr.Server(ctx, budget, 64)
        alone[name] = run(srv, [r], [0])[0]
        srv.close()
    srv = khrr.Server(ctx, budget, 64)
    reqs = [others[0], galaxy, small, sphere, others[1], others[2]]
    mix = run(srv, reqs, [0, 0, 1, 2, 3, 4])
    srv.close()
    for name, k in (("galaxy", 1), ("sphere", 3)):
        a, b = alone[name], mix[k]
        assert a[0] == b[0] and a[1] == b[1], name  # diagnostics, exact
        np.testing.assert_array_equal(a[2], b[2], err_msg=name)
    # every grid holds its request's FP64 mass (bodies outside the box sit
    # in the edge cells)
    for k, r in enumerate(reqs):
        M = r["arrays"][2].sum()
        assert abs(mix[k][2].sum() - M) <= 1e-12 * M, k
        assert mix[k][2].shape == (G,)
    assert abs(mix[1][2].sum() - (1.0 + 6.699356)) < 1e-9
    # the small request's energy against the oracle loop
    pos, vel, mass, eps = small["arrays"]
    h = (0.5 * eps * eps).astype(np.float32).astype(np.float64)
    m32 = mass.astype(np.float32).astype(np.float64)
    _, _, d = orc.run(pos, vel, m32, h, 0.002, small["steps"])
    E = lambda dd: 0.5 * (dd["Ek"] + dd["Ep"])  # noqa: E731
    assert abs(E(mix[2][0]) / (0.5 * (d[0, 0] + d[0, 1])) - 1) < 1e-5
    assert abs(E(mix[2][1]) / (0.5 * (d[-1, 0] + d[-1, 1])) - 1) < 1e-5
