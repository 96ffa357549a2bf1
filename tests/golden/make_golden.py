"""Writes the golden fixtures under tests/golden/ (run once, in the build
container; the outputs are committed, the GPU host never runs this).

  start_S0_head2048.txt  the first 2048 particles of the reference's own
                         input fixture data/nbody/ini/start_S0.txt (data
                         only, header rewritten to the particle count)
  gravity_golden.npz     384 bodies (256 of those stars + 128 Plummer
                         halo bodies with other mass and softening):
                         force and potential evaluated with mpmath at 40
                         significant digits, pairwise, independent of the
                         oracle's FP64 loop
  p2mesh_golden.npz      cell indices of the 2048 stars on the reference's
                         200^3 grid over [-20,20]^3 (nbody.cu:1134-1165)
                         and on a 64^3 grid over [-2,2]^3, from a plain
                         Python loop (IEEE double divide, int() truncation)
  kepler_golden.npz      circular two-body orbit: positions after 200
                         predictor-corrector steps from mpmath evaluation
                         of the reference integrator's recurrence
Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "khrr-nbody_amd"))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def head_fixture():
    src = os.path.join(REF, "data", "nbody", "ini", "start_S0.txt")
    out = os.path.join(HERE, "start_S0_head2048.txt")
    with open(src) as f:
        f.readline()
        lines = [f.readline() for _ in range(2048)]
    with open(out, "w") as f:
        f.write("2048 0.000000\n")
        f.writelines(lines)
    return out


def load_head(path):
    data = np.loadtxt(path, skiprows=1)
    return data[:, :3].T.copy(), data[:, 3:6].T.copy()


def gravity_golden(pos_s):
    import ics
    pd, _ = ics.plummer(128, 6.699356, 0.3333, 1.0, seed=99)
    pos = np.concatenate([pos_s[:, :256], pd], axis=1)
    m = np.concatenate([np.full(256, 1.0 / 262144), np.full(128, 6.699356 / 524288)])
    h = np.concatenate([np.full(256, 0.5 * 0.004 ** 2), np.full(128, 0.5 * 0.006 ** 2)])
    mp.mp.dps = 40
    n = pos.shape[1]
    P = [[mp.mpf(float(pos[c, i])) for i in range(n)] for c in range(3)]
    M = [mp.mpf(float(v)) for v in m]
    H = [mp.mpf(float(v)) for v in h]
    acc = np.zeros((3, n))
    phi = np.zeros(n)
    for i in range(n):
        ax = ay = az = ph = mp.mpf(0)
        for j in range(n):
            dx, dy, dz = P[0][j] - P[0][i], P[1][j] - P[1][i], P[2][j] - P[2][i]
            inv = 1 / mp.sqrt(dx * dx + dy * dy + dz * dz + H[i] + H[j])
            s = M[j] * inv
            s3 = s * inv * inv
            ax += dx * s3
            ay += dy * s3
            az += dz * s3
            ph += s
        acc[:, i] = [float(ax), float(ay), float(az)]
        phi[i] = float(-ph)
    np.savez_compressed(os.path.join(HERE, "gravity_golden.npz"), pos=pos, m=m, h=h, acc=acc,
                        phi=phi)


def cell_of(x, lo, d, s):
    def one(v, lo_, d_, size):
        q = (v - lo_) / d_
        k = 0 if q != q else int(q)  # int() truncates toward zero
        return min(max(k, 0), size - 1)
    ix = one(x[0], lo, d, s)
    iy = one(x[1], lo, d, s)
    iz = one(x[2], lo, d, s)
    return iz * s * s + iy * s + ix


def p2mesh_golden(pos):
    n = pos.shape[1]
    a = [cell_of(pos[:, i].tolist(), -20.0, 0.2, 200) for i in range(n)]
    b = [cell_of(pos[:, i].tolist(), -2.0, 4.0 / 64, 64) for i in range(n)]
    np.savez_compressed(os.path.join(HERE, "p2mesh_golden.npz"), pos=pos,
                        cell_200=np.array(a, np.int32), cell_64=np.array(b, np.int32))


def kepler_golden():
    # two equal masses m = 0.5 on a circular orbit of separation 1 (G = 1):
    # each at radius 0.5, speed v = sqrt(G m_other * r / d^2) = 0.5
    mp.mp.dps = 40
    m = [mp.mpf("0.5"), mp.mpf("0.5")]
    # h = eps^2/2 per body; nonzero so the self pair (d = 0) contributes 0
    # instead of 0 * inf, as in the reference's sum over j = i
    h = mp.mpf(2) ** -34
    x = [[mp.mpf("-0.5"), mp.mpf(0), mp.mpf(0)], [mp.mpf("0.5"), mp.mpf(0), mp.mpf(0)]]
    v = [[mp.mpf(0), mp.mpf("-0.5"), mp.mpf(0)], [mp.mpf(0), mp.mpf("0.5"), mp.mpf(0)]]
    dt = mp.mpf("0.01")

    def acc(p):
        out = []
        for i in range(2):
            j = 1 - i
            d = [p[j][c] - p[i][c] for c in range(3)]
            r2 = sum(t * t for t in d) + 2 * h
            out.append([m[j] * d[c] / r2 ** mp.mpf(1.5) for c in range(3)])
        return out

    a = acc(x)
    for _ in range(200):
        vt = [[v[i][c] + dt * a[i][c] for c in range(3)] for i in range(2)]
        xt = [[x[i][c] + dt / 2 * (v[i][c] + vt[i][c]) for c in range(3)] for i in range(2)]
        at = acc(xt)
        v = [[(v[i][c] + vt[i][c] + dt * at[i][c]) / 2 for c in range(3)] for i in range(2)]
        x = xt
        a = at
    np.savez_compressed(os.path.join(HERE, "kepler_golden.npz"),
                        pos=np.array([[float(x[i][c]) for i in range(2)] for c in range(3)]),
                        vel=np.array([[float(v[i][c]) for i in range(2)] for c in range(3)]),
                        dt=0.01, steps=200)


if __name__ == "__main__":
    path = head_fixture()
    pos_s, _ = load_head(path)
    p2mesh_golden(pos_s)
    kepler_golden()
    gravity_golden(pos_s)
    print("golden fixtures written to", HERE)
