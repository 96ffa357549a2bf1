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
