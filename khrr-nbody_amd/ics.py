"""Seeded synthetic initial conditions for the KHRR-NBody input formats.

The reference ships one star file (data/nbody/ini/start_S0.txt) and no
generator; its DM file is absent.  This module builds the galaxies the
workloads are defined on and writes them in the reference's text format
(`N t` header, then `x y z vx vy vz` per line; read at nbody.cu:1331-1341):

  * stars: exponential disk, radial scale Rd, truncated at Rmax, sech^2
    vertical profile of scale z0; circular speed from the enclosed mass of
    stars + DM, plus Gaussian dispersion of 10 % of v_c per component;
  * dark matter / spheres: Plummer sphere of scale a truncated at r_t,
    isotropic velocities from the Plummer distribution function
    (Aarseth, Henon & Wielen 1974 rejection sampling of q = v/v_esc).

G = 1.  Every generator takes a numpy PCG64 seed, so inputs are
reproducible.  Arrays are float64 (3, n) position / velocity planes.
"""
from __future__ import annotations

import os

import numpy as np


def plummer(n, mass=1.0, a=0.3333, r_t=1.0, seed=0, rng=None):
    """Truncated Plummer sphere of total (truncated) mass `mass`."""
    rng = rng or np.random.default_rng(seed)
    frac = r_t ** 3 / (r_t ** 2 + a ** 2) ** 1.5  # mass fraction inside r_t
    m_full = mass / frac
    u = rng.uniform(0.0, frac, n)
    u = np.maximum(u, 1e-300)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    pos = _iso(r, rng)
    # q = v / v_esc with density g(q) = q^2 (1 - q^2)^(7/2), max 0.092 at q=sqrt(2/9)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x = rng.uniform(0.0, 1.0, todo.size)
        y = rng.uniform(0.0, 0.1, todo.size)
        ok = y < x * x * (1.0 - x * x) ** 3.5
        q[todo[ok]] = x[ok]
        todo = todo[~ok]
    v_esc = np.sqrt(2.0 * m_full / np.sqrt(r * r + a * a))
    vel = _iso(q * v_esc, rng)
    return pos, vel


def _iso(r, rng):
    n = r.shape[0]
    cth = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    sth = np.sqrt(1.0 - cth * cth)
    return np.stack([r * sth * np.cos(ph), r * sth * np.sin(ph), r * cth])


def exp_disk_positions(n, rd=0.25, r_max=1.5, z0=0.025, rng=None):
    """Exponential disk (Sigma ~ exp(-R/Rd), R < r_max), sech^2(z/z0)."""
    grid = np.linspace(0.0, r_max, 20001)
    cdf = 1.0 - (1.0 + grid / rd) * np.exp(-grid / rd)
    cdf /= cdf[-1]
    R = np.interp(rng.uniform(0.0, 1.0, n), cdf, grid)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    z = z0 * np.arctanh(np.clip(rng.uniform(-1.0, 1.0, n), -1 + 1e-12, 1 - 1e-12))
    return np.stack([R * np.cos(ph), R * np.sin(ph), z])


def galaxy(n_s, n_dm, m_s=1.0, m_dm=6.699356, rd=0.25, r_max=1.5, z0=0.025, a_dm=0.3333,
           r_t=1.0, disp=0.1, seed=12345):
    """Disk + halo galaxy of the configurations' recipe.  Returns
    (pos_s, vel_s, pos_dm, vel_dm)."""
    rng = np.random.default_rng(seed)
    pos_dm, vel_dm = plummer(n_dm, m_dm, a_dm, r_t, rng=rng) if n_dm else (np.zeros((3, 0)),) * 2
    pos_s = exp_disk_positions(n_s, rd, r_max, z0, rng)
    # enclosed mass M(<r) of both components at each star's spherical radius
    r_all = np.concatenate([np.linalg.norm(pos_s, axis=0), np.linalg.norm(pos_dm, axis=0)])
    m_all = np.concatenate([np.full(n_s, m_s / max(n_s, 1)), np.full(n_dm, m_dm / max(n_dm, 1))])
    order = np.argsort(r_all, kind="stable")
    m_cum = np.cumsum(m_all[order])
    r_sorted = r_all[order]
    R = np.hypot(pos_s[0], pos_s[1])
    r_s = np.linalg.norm(pos_s, axis=0)
    m_in = np.where(np.searchsorted(r_sorted, r_s, side="left") > 0,
                    m_cum[np.maximum(np.searchsorted(r_sorted, r_s, side="left") - 1, 0)], 0.0)
    vc = np.sqrt(np.maximum(m_in, 0.0) / np.maximum(r_s, 1e-12)) * (R / np.maximum(r_s, 1e-12))
    ph = np.arctan2(pos_s[1], pos_s[0])
    vR = rng.normal(0.0, disp, n_s) * vc
    vphi = vc + rng.normal(0.0, disp, n_s) * vc
    vz = rng.normal(0.0, disp, n_s) * vc
    vel_s = np.stack([vR * np.cos(ph) - vphi * np.sin(ph), vR * np.sin(ph) + vphi * np.cos(ph), vz])
    return pos_s, vel_s, pos_dm, vel_dm


def system_arrays(pos_s, vel_s, pos_dm, vel_dm, m_s, m_dm, eps_s, eps_d):
    """Concatenates stars then DM (the reference's particle order,
    nbody.cu:1318-1417) into pos, vel (float64 planes), mass (float64 per
    particle) and h = eps^2/2."""
    n_s, n_d = pos_s.shape[1], pos_dm.shape[1]
    pos = np.concatenate([pos_s, pos_dm], axis=1)
    vel = np.concatenate([vel_s, vel_dm], axis=1)
    mass = np.concatenate([np.full(n_s, m_s / max(n_s, 1)), np.full(n_d, m_dm / max(n_d, 1))])
    h = np.concatenate([np.full(n_s, 0.5 * eps_s * eps_s), np.full(n_d, 0.5 * eps_d * eps_d)])
    return pos, vel, mass, h


def place(pos, vel, center, vcenter, alpha_deg):
    """A galaxy's rotation about y by alpha and its translation, applied in
    place as the reference's loader does: the z update reads the already
    rotated x (and vz the rotated vx), nbody.cu:1344-1358."""
    ca, sa = np.cos(np.radians(alpha_deg)), np.sin(np.radians(alpha_deg))
    pos = np.array(pos, dtype=np.float64)
    vel = np.array(vel, dtype=np.float64)
    pos[0] = center[0] + pos[0] * ca + pos[2] * sa
    pos[1] += center[1]
    pos[2] = center[2] + pos[2] * ca - pos[0] * sa
    vel[0] = vcenter[0] + vel[0] * ca + vel[2] * sa
    vel[1] += vcenter[1]
    vel[2] = vcenter[2] + vel[2] * ca - vel[0] * sa
    return pos, vel


def encounter(n_s=524288, n_dm=1048576, eps=0.004, seeds=(12346, 12347)):
    """Two disk + halo galaxies at X = +2 (tilt 0, V = (-0.3, +0.05, 0)) and
    X = -2 (tilt 45 degrees, V = (+0.3, -0.05, 0)).  Particle order as the
    reference loads them: stars of both galaxies, then their halos.
    Returns pos, vel, mass, h."""
    setup = (((2.0, 0.0, 0.0), (-0.3, 0.05, 0.0), 0.0), ((-2.0, 0.0, 0.0), (0.3, -0.05, 0.0), 45.0))
    stars, halos = [], []
    for seed, (c, v, a) in zip(seeds, setup):
        ps, vs, pd, vd = galaxy(n_s, n_dm, seed=seed)
        stars.append(place(ps, vs, c, v, a))
        halos.append(place(pd, vd, c, v, a))
    pos = np.concatenate([p for p, _ in stars] + [p for p, _ in halos], axis=1)
    vel = np.concatenate([v for _, v in stars] + [v for _, v in halos], axis=1)
    mass = np.concatenate([np.full(2 * n_s, 1.0 / n_s), np.full(2 * n_dm, 6.699356 / n_dm)])
    h = np.full(pos.shape[1], 0.5 * eps * eps)
    return pos, vel, mass, h


def write_start(path, pos, vel, t=0.0):
    """Reference text format: `N t` then `x y z vx vy vz` lines (6 decimals
    as in the shipped start_S0.txt)."""
    n = pos.shape[1]
    data = np.concatenate([pos, vel]).T
    with open(path, "w") as f:
        f.write(f"{n} {t:.6f}\n")
        np.savetxt(f, data, fmt="%.6f")


def read_start(path):
    """Reads a start_*.txt file; returns (pos, vel, t)."""
    with open(path) as f:
        head = f.readline().split()
        n, t = int(head[0]), float(head[1])
        data = np.loadtxt(f, max_rows=n, ndmin=2)
    return data[:, :3].T.copy(), data[:, 3:6].T.copy(), t


def write_inputs(ini_dir, galaxies, tmax=0.1, dt_save=0.01, gr=None):
    """Writes __start_galaxies.ini, __start_nbody.ini, __gr_par.ini and the
    start files in the reference's layout (parse.cuh:320-367,
    nbody.cu:961-984, 1080-1106).  galaxies: list of dicts with keys
    n_s, n_dm, m_s, m_dm, eps_s, eps_d, alpha, center, vcenter, seed."""
    os.makedirs(ini_dir, exist_ok=True)
    gr = dict(Mh=0.0, a=0.3333, Rh=1.0, Mb=0.0, b=0.0431, Rb=0.2, eps=0.004, dtgrav=0.002,
              K_m=3.72, K_r=0.9) | (gr or {})
    with open(os.path.join(ini_dir, "__start_galaxies.ini"), "w") as f:
        f.write(f"{len(galaxies)}\t\t\t---- M_glx\n")
        for k, g in enumerate(galaxies):
            c, v = g.get("center", (0, 0, 0)), g.get("vcenter", (0, 0, 0))
            f.write(f"{k}\t\t\t---- k_glx\n")
            f.write(f"{g['n_s']},{g['n_dm']} \t\t---- N_s,N_dm\n")
            f.write(f"{g.get('m_s', 1.0)},{g.get('m_dm', 6.699356)} \t\t---- Mass_s,Mass_dm\n")
            f.write(f"{g.get('eps_s', 0.004)},{g.get('eps_d', 0.004)} \t\t---- eps_s,eps_dm\n")
            f.write(f"{g.get('alpha', 0.0)}\t\t\t---- alpha\n")
            f.write(f"{c[0]},{c[1]},{c[2]}\t\t---- X,Y,Z\n")
            f.write(f"{v[0]},{v[1]},{v[2]}\t\t---- Vx,Vy,Vz\n")
    with open(os.path.join(ini_dir, "__start_nbody.ini"), "w") as f:
        f.write(f"0 \t--- i_cont\n{tmax} \t--- tmax\n{dt_save} \t--- dt_save\n")
    with open(os.path.join(ini_dir, "__gr_par.ini"), "w") as f:
        for key in ("Mh", "a", "Rh", "Mb", "b", "Rb", "eps", "dtgrav", "K_m", "K_r"):
            f.write(f"{gr[key]} \t--- {key}\n")
    for k, g in enumerate(galaxies):
        ps, vs, pd, vd = galaxy(g["n_s"], g["n_dm"], g.get("m_s", 1.0), g.get("m_dm", 6.699356),
                                seed=g.get("seed", 12345 + k))
        if g["n_s"] and g.get("m_s", 1.0) > 0:
            write_start(os.path.join(ini_dir, f"start_S{k}.txt"), ps, vs)
        if g["n_dm"] and g.get("m_dm", 6.699356) > 0:
            write_start(os.path.join(ini_dir, f"start_DM{k}.txt"), pd, vd)


def plummer_population(n_sys=512, n_min=1024, n_max=131072, seed=7):
    """Sizes of the batched workload: log-uniform on [n_min, n_max] rounded
    to a multiple of 1024; softening eps = 0.004 (N/16384)^(-1/3)."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_sys))
    n = np.clip(np.round(n / 1024.0) * 1024, n_min, n_max).astype(np.int64)
    eps = 0.004 * (n / 16384.0) ** (-1.0 / 3.0)
    return n, eps
