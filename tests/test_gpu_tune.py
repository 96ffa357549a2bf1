"""The project's autotuner (khrr_tune_gravity, khrr_tune_batched) and its
JSON cache (khrr_tune_save / khrr_tune_load).

* every shape the tuner may pick -- one-sided j-tiles 128..1024,
  pair-symmetric i_per_lane 4/8 x superblocks 1024/2048/4096 -- and the
  shapes it picked are within the force bar against the FP64 oracle
  (ACCEL/PHI, nbody-kernel.cuh:43-142): median <= 1e-6, p99 <= 1e-5;
* the batched form with a tuned j-tile equals khrr_accel with the same
  shape bit for bit, system by system;
* the cache round-trips through a file into a fresh context, and a file
  from another device, a malformed file or an invalid shape is refused
  with nothing loaded.

Each test works on a context of its own, so the tuned shapes it stores do
not reach the other tests' session context."""
import json

import numpy as np
import pytest

import ics
import oracle as orc
from helpers import err_stats, rel_err, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture()
def own_ctx():
    import khrr
    c = khrr.Context(0)
    yield c
    c.close()


def _plummer(n, seed):
    pos, _ = ics.plummer(n, 1.0, 0.3333, 1.0, seed=seed)
    eps = 0.004 * (n / 16384.0) ** (-1.0 / 3.0)
    return (pos.astype(np.float32), np.full(n, 1.0 / n, np.float32),
            np.full(n, 0.5 * eps * eps, np.float32))


def _accel(ctx, p32, m32, h32, tune):
    import torch
    n = p32.shape[1]
    p, m, h = to_dev(p32, torch.float32), to_dev(m32, torch.float32), to_dev(h32, torch.float32)
    acc = torch.zeros((3, n), device="cuda:0")
    ctx.accel(p, h, p, m, h, acc=acc, tune=tune)
    torch.cuda.synchronize()
    return acc.cpu().numpy()


def _check_rows(acc, p32, m32, h32, rows, what):
    ref, _ = orc.accel_rows(rows, p32.astype(np.float64), h32.astype(np.float64),
                            m32.astype(np.float64))
    st = err_stats(rel_err(acc[:, rows].astype(np.float64), ref))
    assert st["median"] <= 1e-6 and st["p99"] <= 1e-5, (what, st)
    return st


def test_every_candidate_shape_within_force_bar(own_ctx):
    import khrr
    n = 40000  # superblocks of 1024, 2048 and 4096 all give two or more
    p32, m32, h32 = _plummer(n, 4242)
    rows = np.sort(np.random.default_rng(5).choice(n, 2048, replace=False)).astype(np.int32)
    shapes = [khrr.Tune(2, 4, 4, jt, 0, 0) for jt in (128, 256, 512, 1024)]
    shapes += [khrr.Tune(ipl, 4, 1, 256, 1, sb)
               for ipl, sb in ((4, 1024), (4, 2048), (8, 2048), (4, 4096), (8, 4096))]
    for t in shapes:
        acc = _accel(own_ctx, p32, m32, h32, t)
        st = _check_rows(acc, p32, m32, h32, rows, t.as_dict())
        print("shape", t.as_dict(), st)


@pytest.mark.parametrize("n", [3000, 40000])
def test_tuned_shape_within_force_bar(own_ctx, n):
    t, ms = own_ctx.tune_gravity(n)
    print("tuned", n, t.as_dict(), f"{ms:.3f} ms")
    cached = own_ctx.tune_cached(n)
    assert cached is not None and cached.as_dict() == t.as_dict()
    assert t.j_tile in (128, 256, 512, 1024)
    assert t.superblock in (0, 1024, 2048, 4096)
    p32, m32, h32 = _plummer(n, 77 + n)
    acc = _accel(own_ctx, p32, m32, h32, None)  # tune NULL: the cached shape
    np.testing.assert_array_equal(acc, _accel(own_ctx, p32, m32, h32, t))
    rows = np.sort(np.random.default_rng(n).choice(n, min(n, 2048), replace=False))
    _check_rows(acc, p32, m32, h32, rows.astype(np.int32), t.as_dict())


@pytest.mark.parametrize("j_tile", [128, 512, 1024])
def test_batched_j_tile_equals_plain(own_ctx, j_tile):
    import khrr
    import torch
    sizes = [3000, 7168, 20000, 1024, 512]
    parts = [_plummer(n, 900 + k) for k, n in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    pos = np.concatenate([p[0] for p in parts], axis=1)
    m = np.concatenate([p[1] for p in parts])
    h = np.concatenate([p[2] for p in parts])
    N = int(off[-1])
    acc = torch.zeros((3, N), device="cuda:0")
    own_ctx.accel_batched(to_dev(off, torch.int32), off.tolist(), to_dev(pos, torch.float32),
                          to_dev(h, torch.float32), to_dev(m, torch.float32), acc=acc,
                          tune=khrr.Tune(4, 2, 1, j_tile, 0, 0))
    torch.cuda.synchronize()
    a = acc.cpu().numpy()
    for k, n in enumerate(sizes):
        p32, m32, h32 = parts[k]
        if n >= 16384:  # pair-symmetric path: the rule's superblock
            plain = khrr.Tune(8, 4, 1, 256, 1, 0)
        else:
            plain = khrr.Tune(4, 4, max(1, min(32, -(-n // 1024))), j_tile, 0, 0)
        np.testing.assert_array_equal(a[:, off[k]:off[k + 1]],
                                      _accel(own_ctx, p32, m32, h32, plain),
                                      err_msg=f"system {k} (n={n}), j_tile {j_tile}")


def test_cache_round_trip_and_refusals(own_ctx, tmp_path):
    import khrr
    sizes = [4096, 1024, 20000]
    tg, _ = own_ctx.tune_gravity(5000)
    tb, _ = own_ctx.tune_batched(np.concatenate([[0], np.cumsum(sizes)]).tolist())
    path = str(tmp_path / "tune.json")
    own_ctx.tune_save(path)
    doc = json.load(open(path))
    assert doc["format"] == "khrr-tune" and doc["version"] == 1 and doc["cu_count"] > 0
    assert [e["bucket"] for e in doc["gravity"]] == [13]  # 5000 -> ceil(log2) = 13
    fresh = khrr.Context(0)
    try:
        assert fresh.tune_cached(5000) is None and fresh.tune_batched_cached() is None
        fresh.tune_load(path)
        assert fresh.tune_cached(5000).as_dict() == tg.as_dict()
        assert fresh.tune_cached(4097).as_dict() == tg.as_dict()  # same size bucket
        assert fresh.tune_cached(4096) is None
        assert fresh.tune_batched_cached().as_dict() == tb.as_dict()
    finally:
        fresh.close()

    def refused(mutate, code, match):
        d = json.load(open(path))
        mutate(d)
        bad = str(tmp_path / "bad.json")
        with open(bad, "w") as f:
            f.write(d if isinstance(d, str) else json.dumps(d))
        c = khrr.Context(0)
        try:
            with pytest.raises(khrr.KhrrError, match=match) as ei:
                c.tune_load(bad)
            assert ei.value.code == code
            assert c.tune_cached(5000) is None and c.tune_batched_cached() is None
        finally:
            c.close()

    refused(lambda d: d.update(arch="gfx000"), -4, "tuned on gfx000")
    refused(lambda d: d.update(cu_count=d["cu_count"] + 1), -4, "CUs")
    refused(lambda d: d["gravity"][0].update(j_tile=100), -1, "j_tile")
    refused(lambda d: d["gravity"][0].update(superblock=3000), -1, "superblock")
    refused(lambda d: d["gravity"][0].pop("i_per_lane"), -1, "i_per_lane")
    refused(lambda d: d.update(version=2), -1, "version 1")
    with open(str(tmp_path / "trunc.json"), "w") as f:
        f.write(open(path).read()[:-10])
    c = khrr.Context(0)
    try:
        with pytest.raises(khrr.KhrrError, match="not a JSON object"):
            c.tune_load(str(tmp_path / "trunc.json"))
        with pytest.raises(khrr.KhrrError, match="can't open"):
            c.tune_load(str(tmp_path / "missing.json"))
    finally:
        c.close()
