"""Text of the driver's debug dumps and cell-potential slices, on the CPU.

The driver's writers (khrr-nbody_amd/host/nbody_io.cpp) are compiled with a
small harness (tests/host_formats.cpp) and their files compared line by line
with the text the reference's writers produce for the same values:
std::format "{}" of the x coordinate (the shortest form that reads back to
the same double, no trailing ".0"), "{:.7f}" / "{:.6f}" values
(print_cell_phi nbody.cu:704-724, check_mass :107-112, the slices of
convert_particles :412-645), "i, acc" rows (:622-697), counts through
std::ostream's default format (%g, :412-432) and the count distribution with
its running sum (check_particles_distribution :68-96)."""
import os
import shutil
import subprocess

import pytest

from conftest import PKG

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dumps(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("formats")
    exe = str(d / "host_formats")
    host = os.path.join(PKG, "host")
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", host,
                    os.path.join(ROOT, "tests", "host_formats.cpp"),
                    os.path.join(host, "nbody_io.cpp"), "-o", exe], check=True)
    subprocess.run([exe, str(d)], check=True)
    return d


def _shortest(x):
    # std::format "{}" of a double: shortest round-trip digits, as repr,
    # without repr's ".0" on integral values
    r = repr(float(x))
    return r[:-2] if r.endswith(".0") else r


def _inputs():
    nx = 5
    cells = [1.0 / (c + 3) - 0.01 * c for c in range(nx ** 3)]
    count = [(c * 7) % 5 + (1234567 if c == 62 else 0) for c in range(nx ** 3)]
    return nx, 0.5 * nx * 0.2, 0.2, cells, count


def _lines(d, name):
    return open(os.path.join(d, name)).read().split("\n")


def test_x_slices(dumps):
    nx, sim_l, dx, cells, _ = _inputs()
    line = [i + (nx // 2) * nx + (nx // 2) * nx * nx for i in range(nx)]
    for name, header, dec in (("slice7.txt", "x, phi", 7), ("slice6.txt", "x, mass", 6)):
        want = [header] + [f"{_shortest(-sim_l + i * dx)}, {cells[line[i]]:.{dec}f}"
                           for i in range(nx)] + [""]
        assert _lines(dumps, name) == want
    # the x column exercises a non-integral sum and an exact zero
    assert _lines(dumps, "slice7.txt")[1].startswith("-0.5, ")
    assert _lines(dumps, "slice7.txt")[4].startswith("0.10000000000000009, ")


def test_index_list(dumps):
    body = [0.1, -2.5, 3.14159265358979, 1e-9, 12345.678901234567, 0.0]
    want = ["i, acc"] + [f"{i}, {v:.7f}" for i, v in enumerate(body)] + [""]
    assert _lines(dumps, "list.txt") == want


def test_count_slice(dumps):
    nx, _, _, _, count = _inputs()
    row = (nx // 2) * nx + (nx // 2) * nx * nx
    want = ["xi, counts"] + [f"{i}, {format(float(count[row + i]), 'g')}"
                             for i in range(nx)] + [""]
    assert _lines(dumps, "counts.txt") == want
    assert "1.23457e+06" in _lines(dumps, "counts.txt")[3]


def test_count_distribution(dumps):
    _, _, _, _, count = _inputs()
    dist = {}
    for c in count:
        dist[c] = dist.get(c, 0) + 1
    rows = sorted(dist.items(), reverse=True)
    want = ["n_particles, cells_with_so_many_particles"] + [f"{k}, {v}" for k, v in rows] + [""]
    assert _lines(dumps, "particles_in_cell.txt") == want
    partial, run = [], 0.0
    for k, v in rows:
        run += float(k) * v
        partial.append(run)
    want = [f"{i}, {format(partial[i], 'g')}" for i in range(1, len(partial))] + [""]
    assert _lines(dumps, "particles_in_cell_partial_sum.txt") == want
