"""The drop-in driver bin/khrr_nbody (replaces run() of nbody/src/nbody.cu)
on inputs in the reference's formats.  CPU tests cover the input errors;
the GPU test runs a small galaxy and checks the written files: LIE_0.bin
layout, snapshot layout and names, CSV headers, and the conservation series
against the FP64 oracle loop (tolerance: relative energy difference
<= 1e-5 of |E0| per save, FP32 forces vs FP64)."""
import os
import struct
import subprocess

import numpy as np
import pytest

import ics
from conftest import PKG

BIN = os.path.join(PKG, "bin", "khrr_nbody")


def _layout(tmp, galaxies, tmax, dtsave, gr=None):
    base = os.path.join(tmp, "data", "nbody")
    for d in ("ini", "bin", "out"):
        os.makedirs(os.path.join(base, d), exist_ok=True)
    ics.write_inputs(os.path.join(base, "ini"), galaxies, tmax=tmax, dt_save=dtsave, gr=gr)
    return base


def test_missing_inputs_are_reported(tmp_path):
    r = subprocess.run([BIN, "--data", str(tmp_path / "nowhere")], capture_output=True,
                       text=True)
    assert r.returncode == 1
    assert "galaxy properties" in r.stderr


def test_missing_start_file_is_reported(tmp_path):
    base = _layout(str(tmp_path), [dict(n_s=64, n_dm=64, seed=1)], 0.004, 0.002)
    os.remove(os.path.join(base, "ini", "start_DM0.txt"))
    r = subprocess.run([BIN, "--data", base], capture_output=True, text=True)
    assert r.returncode == 1
    assert "start_DM0.txt" in r.stderr


def test_bad_arguments():
    r = subprocess.run([BIN, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 2


def test_written_inputs_parse_like_reference(tmp_path):
    # the generator's files follow the reference's fscanf patterns
    base = _layout(str(tmp_path), [dict(n_s=32, n_dm=16, seed=3, alpha=45.0,
                                        center=(2.0, 0.0, 0.0), vcenter=(-0.3, 0.05, 0.0))],
                   0.1, 0.01)
    txt = open(os.path.join(base, "ini", "__start_galaxies.ini")).read().splitlines()
    assert txt[0].split()[0] == "1" and txt[2].split()[0] == "32,16"
    pos, vel, t = ics.read_start(os.path.join(base, "ini", "start_S0.txt"))
    assert pos.shape == (3, 32) and t == 0.0


def _oracle_series(base, steps):
    """FP64 oracle of the same run, diagnostics after every step."""
    import oracle as orc
    ps, vs, _ = ics.read_start(os.path.join(base, "ini", "start_S0.txt"))
    pd, vd, _ = ics.read_start(os.path.join(base, "ini", "start_DM0.txt"))
    pos = np.concatenate([ps, pd], axis=1)
    vel = np.concatenate([vs, vd], axis=1)
    n_s, n_d = ps.shape[1], pd.shape[1]
    mass = np.concatenate([np.full(n_s, 1.0 / n_s), np.full(n_d, 6.699356 / n_d)])
    h = np.full(n_s + n_d, 0.5 * 0.004 ** 2).astype(np.float32).astype(np.float64)
    _, _, d = orc.run(pos, vel, mass, h, 0.002, steps)
    return d


@pytest.mark.gpu
def test_small_galaxy_run_outputs(tmp_path):
    base = _layout(str(tmp_path), [dict(n_s=1024, n_dm=2048, seed=5)], 0.02, 0.002)
    r = subprocess.run([BIN, "--data", base, "--grid", "32", "0.125", "--json"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out, bin_ = os.path.join(base, "out"), os.path.join(base, "bin")
    # LIE_0.bin: L0, P0, E0 + 16 bytes (nbody.cu:844-848)
    raw = open(os.path.join(bin_, "LIE_0.bin"), "rb").read()
    assert len(raw) == 72
    L0 = struct.unpack("3d", raw[:24])
    E0 = struct.unpack("d", raw[48:56])[0]
    # 10 steps of dt 0.002 with dt_save 0.002: saves 1..10, snapshots at 10
    for name, count in (("S", 1024), ("DM", 2048)):
        path = os.path.join(bin_, f"{name}_   10.bin")
        data = open(path, "rb").read()
        assert len(data) == 4 + 8 + 48 * count
        assert struct.unpack("i", data[:4])[0] == count
        assert abs(struct.unpack("d", data[4:12])[0] - 0.02) < 1e-12
    E = np.genfromtxt(os.path.join(out, "E(t).csv"), delimiter=",", skip_header=1)
    assert E.shape == (11, 7)
    assert open(os.path.join(out, "E(t).csv")).readline().strip() == \
        "it, t, E, E_rel_err, E_rel_err_magnitude, Ek, Ep"
    for f in ("Imp(t).csv", "LL(t).csv", "Lz(t).csv", "sample.csv"):
        assert os.path.exists(os.path.join(out, f))
    assert os.path.exists(os.path.join(out, "phi", "phi_   10.csv"))
    d = _oracle_series(base, 10)
    E_ref = 0.5 * (d[:, 0] + d[:, 1])
    assert abs(E0 - E_ref[0]) <= 1e-5 * abs(E_ref[0])
    np.testing.assert_allclose(E[:, 2], E_ref, rtol=1e-5)
    np.testing.assert_allclose(L0[2], d[0, 7], rtol=1e-9)
    np.testing.assert_array_equal(E[:, 0], np.arange(11))


@pytest.mark.gpu
def test_wave_solver_run(tmp_path):
    # the reference's NBodySolver::Wave branch: mesh force from the damped
    # wave field (setup iterations, then dtgrav/dt_wave per step)
    base = _layout(str(tmp_path), [dict(n_s=1024, n_dm=2048, seed=6)], 0.02, 0.002)
    r = subprocess.run([BIN, "--data", base, "--grid", "32", "0.125", "--solver", "wave",
                        "--setup-iters", "600", "--max-steps", "4", "--json"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = os.path.join(base, "out")
    E = np.genfromtxt(os.path.join(out, "E(t).csv"), delimiter=",", skip_header=1)
    assert E.shape == (5, 7) and np.isfinite(E).all()
    # Ep = sum m phi with the mesh potential: bound system, negative
    assert (E[:, 6] < 0).all()
    for frame in ("    0", "    1", "    4"):
        assert os.path.exists(os.path.join(out, "phi", f"phi_{frame}.csv"))
    assert os.path.getsize(os.path.join(base, "bin", "LIE_0.bin")) == 72


def _rows(path, header):
    lines = open(path).read().splitlines()
    assert lines[0] == header, (path, lines[0])
    return np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])


@pytest.mark.gpu
def test_debug_dumps(tmp_path):
    # convert_particles (nbody.cu:287-701) with its checks (:35-127): the
    # files of data/nbody/debug after the initial forces, against the FP64
    # oracle of the same chain.  Integer files exactly; the binning-only
    # values to their printed decimals; values through the FP32 forces
    # within 1e-4, and through the FP32 wave field within 1e-4 of their
    # largest value (the stencil's own tolerance is 1e-5 per field).
    import oracle as orc
    base = _layout(str(tmp_path), [dict(n_s=1024, n_dm=2048, seed=7)], 0.02, 0.002)
    nx, dx, iters = 16, 0.25, 50
    r = subprocess.run([BIN, "--data", base, "--grid", str(nx), str(dx), "--debug-dumps",
                        "--setup-iters", str(iters), "--max-steps", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    dbg = os.path.join(base, "debug")
    ps, _, _ = ics.read_start(os.path.join(base, "ini", "start_S0.txt"))
    pd, _, _ = ics.read_start(os.path.join(base, "ini", "start_DM0.txt"))
    pos = np.concatenate([ps, pd], axis=1)
    n_s, n_d = ps.shape[1], pd.shape[1]
    mass = np.concatenate([np.full(n_s, 1.0 / n_s), np.full(n_d, 6.699356 / n_d)])
    h = np.full(n_s + n_d, 0.5 * 0.004 ** 2).astype(np.float32).astype(np.float64)
    b = orc.p2mesh_bin([-2.0] * 3, [dx] * 3, [nx] * 3, pos, mass)
    G = nx ** 3
    line = np.arange(nx) + (nx // 2) * nx + (nx // 2) * nx * nx
    count, cm, cid = b["count"], b["cell_mass"], b["cell_id"]
    # check_particles_distribution
    vals, freq = np.unique(count, return_counts=True)
    d = _rows(os.path.join(dbg, "particles_in_cell.txt"),
              "n_particles, cells_with_so_many_particles")
    np.testing.assert_array_equal(d, np.stack([vals[::-1], freq[::-1]], axis=1))
    part = np.cumsum(vals[::-1] * freq[::-1].astype(np.float64))
    ps_rows = np.array([[float(x) for x in ln.split(",")] for ln in
                        open(os.path.join(dbg, "particles_in_cell_partial_sum.txt"))])
    np.testing.assert_array_equal(ps_rows[:, 0], np.arange(1, len(part)))
    np.testing.assert_allclose(ps_rows[:, 1], part[1:], rtol=1e-5)
    # binning slices and per-body mass lists
    x = -2.0 + np.arange(nx) * dx
    d = _rows(os.path.join(dbg, "particle_counts.txt"), "xi, counts")
    np.testing.assert_array_equal(d, np.stack([np.arange(nx), count[line]], axis=1))
    d = _rows(os.path.join(dbg, "mass_cell.txt"), "x, mass")
    np.testing.assert_array_equal(d[:, 0], x)
    np.testing.assert_allclose(d[:, 1], cm[line], atol=5.1e-7)
    d = _rows(os.path.join(dbg, "mpart.txt"), "i, acc")
    np.testing.assert_array_equal(d[:, 0], np.arange(pos.shape[1]))
    np.testing.assert_allclose(d[:, 1], mass, atol=5.1e-8)
    d = _rows(os.path.join(dbg, "mpart_from_cell.txt"), "i, acc")
    np.testing.assert_allclose(d[:, 1], orc.p2mesh_cell_avg(cid, count, cm), atol=5.1e-8)
    # body forces
    acc, phi = orc.accel(pos, h, pos, mass, h)
    pabs = orc.acc_abs(acc)
    d = _rows(os.path.join(dbg, "nbody_part_acc.txt"), "i, acc")
    np.testing.assert_allclose(d[:, 1], pabs, rtol=1e-4, atol=5.1e-8)
    d = _rows(os.path.join(dbg, "nbody_cell_acc.txt"), "x, acc")
    np.testing.assert_allclose(d[:, 1], orc.p2mesh_cell_weighted(cid, cm, mass, pabs)[line],
                               rtol=1e-4, atol=5.1e-8)
    cphi = orc.p2mesh_cell_phi(cid, count, phi)
    d = _rows(os.path.join(dbg, "phi_part.txt"), "x, phi")
    np.testing.assert_allclose(d[:, 1], cphi[line], rtol=1e-4, atol=5.1e-8)
    _, cabs = orc.mesh_acc(nx, dx, cphi)
    d = _rows(os.path.join(dbg, "acc_part.txt"), "x, acc")
    assert np.max(np.abs(d[:, 1] - cabs[line])) <= 1e-4 * np.max(cabs) + 5.1e-8
    # wave field from rest and its acceleration
    wphi, _ = orc.wave_iterate(orc.reference_wave(nx, dx), np.zeros(G), np.zeros(G), cm, iters)
    wacc, wabs = orc.mesh_acc(nx, dx, wphi)
    for name, ref in (("phi_cell.txt", wphi), ("acc_cell.txt", wabs)):
        d = _rows(os.path.join(dbg, name), "x, " + ("phi" if "phi" in name else "acc"))
        assert np.max(np.abs(ref)) > 0
        assert np.max(np.abs(d[:, 1] - ref[line])) <= 1e-4 * np.max(np.abs(ref)) + 5.1e-8, name
    cpabs = orc.acc_abs(wacc[:, cid])
    d = _rows(os.path.join(dbg, "cell_part_acc.txt"), "i, acc")
    assert np.max(np.abs(d[:, 1] - cpabs)) <= 1e-4 * np.max(cpabs) + 5.1e-8
    cyc = orc.p2mesh_cell_weighted(cid, cm, mass, cpabs)
    d = _rows(os.path.join(dbg, "acc_cell_cycle.txt"), "x, acc")
    assert np.max(np.abs(d[:, 1] - cyc[line])) <= 1e-4 * np.max(cpabs) + 5.1e-8
    # console lines of the checks
    out = {}
    for ln in r.stdout.splitlines():
        if ": " in ln:
            k, v = ln.rsplit(": ", 1)
            out[k] = v
    assert "cell info mismatch!!! " not in r.stdout
    ep = np.sum(0.5 * cm * wphi)
    assert abs(float(out["ep"]) - ep) <= 1e-4 * np.sum(0.5 * cm * np.abs(wphi))
    inside = cm.sum()
    np.testing.assert_allclose(float(out["total mass in cells"]), inside, rtol=1e-5)
    np.testing.assert_allclose(float(out["total mass in particles"]), mass.sum(), rtol=1e-5)
    np.testing.assert_allclose(float(out["total mpart from cell"]),
                               np.sum(orc.p2mesh_cell_avg(cid, count, cm)), rtol=1e-5)
    np.testing.assert_allclose(float(out["total mpart"]), mass.sum(), rtol=1e-5)
    np.testing.assert_allclose(float(out["nbody part acc sum"]), pabs.sum(), rtol=1e-4)
    np.testing.assert_allclose(float(out["cell part acc sum"]), cpabs.sum(), rtol=1e-4)


# ---- restart (i_cont > 0), nbody.cu:1247-1314 ----------------------------

def _set_start(base, i_cont, tmax, dtsave):
    """__start_nbody.ini in the reference's layout (nbody.cu:961-984)."""
    with open(os.path.join(base, "ini", "__start_nbody.ini"), "w") as f:
        f.write(f"{i_cont} \t--- i_cont\n{tmax} \t--- tmax\n{dtsave} \t--- dt_save\n")


def _snapshot(path, count, t, records=None):
    """S_/DM_ .bin layout (nbody.cu:726-755): int count, double t, then
    x, y, z, vx, vy, vz per body."""
    records = count if records is None else records
    vals = np.random.default_rng(abs(count) + 1).normal(size=6 * records)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", count) + struct.pack("<d", t) + vals.astype("<f8").tobytes())


@pytest.mark.parametrize("case,message", [
    ("no_stars", "S_   10.bin"),
    ("short_stars", "short snapshot"),
    ("wrong_count", "holds 63 bodies, the galaxies have 64"),
    ("short_header", "short snapshot header"),
    ("no_dm", "DM_   10.bin"),
    ("no_lie", "LIE_0.bin"),
    ("short_lie", "short LIE_0.bin"),
])
def test_restart_input_errors(tmp_path, case, message):
    # every failure is found while reading, before the device is touched
    base = _layout(str(tmp_path), [dict(n_s=64, n_dm=32, seed=9)], 0.04, 0.002)
    _set_start(base, 10, 0.04, 0.002)
    b = os.path.join(base, "bin")
    s_path, d_path = os.path.join(b, "S_   10.bin"), os.path.join(b, "DM_   10.bin")
    if case != "no_stars":
        if case == "short_stars":
            _snapshot(s_path, 64, 0.02, records=10)
        elif case == "wrong_count":
            _snapshot(s_path, 63, 0.02)
        elif case == "short_header":
            open(s_path, "wb").write(struct.pack("<i", 64))
        else:
            _snapshot(s_path, 64, 0.02)
    if case not in ("no_stars", "short_stars", "wrong_count", "short_header", "no_dm"):
        _snapshot(d_path, 32, 0.02)
    if case == "short_lie":
        open(os.path.join(b, "LIE_0.bin"), "wb").write(b"\0" * 40)
    r = subprocess.run([BIN, "--data", base], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, (case, r.stdout[-500:], r.stderr[-500:])
    assert message in r.stderr, (case, r.stderr)


@pytest.mark.gpu
def test_restart_continues_bit_for_bit(tmp_path):
    # 20 saves straight through against 10 saves then a restart from
    # i_cont = 10 to save 20: the snapshots, LIE_0.bin and every CSV row
    # after the restart are byte-identical (the restart recomputes acc from
    # the snapshot's positions, which are the positions the straight run's
    # last force was evaluated at; --clear-cell-phi because the reference's
    # never-cleared cell-potential buffer does not survive a restart)
    gal = [dict(n_s=1024, n_dm=2048, seed=11)]
    a = _layout(str(tmp_path / "a"), gal, 0.04, 0.002)
    b = _layout(str(tmp_path / "b"), gal, 0.02, 0.002)
    args = ["--grid", "32", "0.125", "--clear-cell-phi"]
    for base in (a, b):
        r = subprocess.run([BIN, "--data", base] + args, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
    assert os.path.exists(os.path.join(b, "bin", "S_   10.bin"))
    assert not os.path.exists(os.path.join(b, "bin", "S_   20.bin"))
    _set_start(b, 10, 0.04, 0.002)
    r = subprocess.run([BIN, "--data", b] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    files = ["bin/S_   10.bin", "bin/DM_   10.bin", "bin/S_   20.bin", "bin/DM_   20.bin",
             "bin/LIE_0.bin", "out/E(t).csv", "out/Imp(t).csv", "out/LL(t).csv", "out/Lz(t).csv",
             "out/sample.csv", "out/phi/phi_   11.csv", "out/phi/phi_   20.csv"]
    for f in files:
        fa = open(os.path.join(a, f), "rb").read()
        fb = open(os.path.join(b, f), "rb").read()
        assert fa == fb, f
    E = np.genfromtxt(os.path.join(b, "out", "E(t).csv"), delimiter=",", skip_header=1)
    np.testing.assert_array_equal(E[:, 0], np.arange(21))
    raw = open(os.path.join(b, "bin", "S_   20.bin"), "rb").read()
    assert struct.unpack("<i", raw[:4])[0] == 1024
    assert abs(struct.unpack("<d", raw[4:12])[0] - 0.04) < 1e-12


@pytest.mark.gpu
def test_tune_cache_flag(tmp_path):
    # --tune-cache loads a khrr_tune_save file into the driver's context; a
    # file made on another kind of device is refused before any step
    import json

    import khrr
    c = khrr.Context(0)
    try:
        c.tune_gravity(3072)
        good = str(tmp_path / "tune.json")
        c.tune_save(good)
    finally:
        c.close()
    base = _layout(str(tmp_path), [dict(n_s=1024, n_dm=2048, seed=13)], 0.004, 0.002)
    r = subprocess.run([BIN, "--data", base, "--grid", "16", "0.25", "--tune-cache", good],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    doc = json.load(open(good))
    doc["arch"] = "gfx000"
    bad = str(tmp_path / "other.json")
    json.dump(doc, open(bad, "w"))
    r = subprocess.run([BIN, "--data", base, "--grid", "16", "0.25", "--tune-cache", bad],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "tuned on gfx000" in r.stderr, r.stderr[-2000:]
