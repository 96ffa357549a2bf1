"""ctypes wrapper of the CPU FP64 oracle (oracle/libkhrr_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker / baseline, never by the product.
Arrays are numpy float64 (int32 for integer outputs); "vec3 planes" are
arrays of shape (3, n).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkhrr_oracle.so")

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p


class OrcExtPot(ctypes.Structure):
    _fields_ = [(n, _D) for n in ("Mh", "Mh_inf", "a", "Rh", "Mb", "b", "Rb", "Rh2", "con",
                                  "const1", "c_phi_h", "c_phi_b")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"oracle not built at {LIB_PATH}; run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_extpot_init.argtypes = [_D] * 6 + [ctypes.POINTER(OrcExtPot)]
        L.orc_dev_fex.argtypes = [ctypes.POINTER(OrcExtPot), _P, _P]
        L.orc_accel.argtypes = [_I, _P, _P, _I, _P, _P, _P, _P, _P, _I]
        L.orc_accel_rows.argtypes = [_I, _P, _I, _P, _P, _I, _P, _P, _P, _P, _P]
        L.orc_integrate.argtypes = [_I, _I, _D, _D, ctypes.POINTER(OrcExtPot), _P, _P, _P, _P,
                                    _P, _P]
        L.orc_diag.argtypes = [_I, _P, _P, _P, _P, _P]
        L.orc_run.argtypes = [_I, _P, _P, _P, _P, _D, _I, ctypes.POINTER(OrcExtPot), _I, _P, _I]
        L.orc_p2mesh_bin.argtypes = [_P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]
        L.orc_p2mesh_bin.restype = _I
        L.orc_p2mesh_cell_phi.argtypes = [_I, ctypes.c_longlong, _P, _P, _P, _P, _I]
        L.orc_p2mesh_cell_weighted.argtypes = [_I, ctypes.c_longlong, _P, _P, _P, _P, _P]
        L.orc_p2mesh_cell_avg.argtypes = [_I, _P, _P, _P, _P]
        L.orc_acc_abs.argtypes = [_I, _P, _P]
        L.orc_sat_trunc.argtypes = [_D]
        L.orc_sat_trunc.restype = _I
        L.orc_wave_step.argtypes = [_I] + [_D] * 7 + [_P] * 4
        L.orc_wave_iterate.argtypes = [_I] + [_D] * 7 + [_P, _P, _P, _I, _P, _P]
        L.orc_mesh_acc.argtypes = [_I, _D, _P, _P, _P]
        L.orc_mesh_to_particles.argtypes = [_I, ctypes.c_longlong, _P, _P, _P, _P, _P]
        L.orc_num_threads.restype = _I
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def extpot(Mh, a, Rh, Mb, b, Rb) -> OrcExtPot:
    e = OrcExtPot()
    lib().orc_extpot_init(Mh, a, Rh, Mb, b, Rb, ctypes.byref(e))
    return e


def dev_fex(e: OrcExtPot, p):
    pp = _f64(p)
    f = np.zeros(3)
    lib().orc_dev_fex(ctypes.byref(e), _p(pp), _p(f))
    return f


def accel(pos_i, h_i, pos_j=None, m_j=None, h_j=None, acc=None, phi=None, accumulate=False,
          want_acc=True, want_phi=True):
    """ACCEL_kernel + PHI_kernel in FP64.  Returns (acc (3,n_i), phi (n_i,))."""
    pos_i = _f64(pos_i)
    h_i = _f64(h_i)
    if pos_j is None:
        pos_j, h_j = pos_i, h_i
    pos_j, m_j, h_j = _f64(pos_j), _f64(m_j), _f64(h_j)
    n_i, n_j = pos_i.shape[1], pos_j.shape[1]
    if want_acc:
        acc = np.zeros((3, n_i)) if acc is None else _f64(acc).copy()
    else:
        acc = None
    if want_phi:
        phi = np.zeros(n_i) if phi is None else _f64(phi).copy()
    else:
        phi = None
    lib().orc_accel(n_i, _p(pos_i), _p(h_i), n_j, _p(pos_j), _p(m_j), _p(h_j), _p(acc), _p(phi),
                    1 if accumulate else 0)
    return acc, phi


def accel_rows(rows, pos, h, m):
    """FP64 force and potential for the listed rows of one system."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    pos, h, m = _f64(pos), _f64(h), _f64(m)
    n = pos.shape[1]
    acc = np.zeros((3, len(rows)))
    phi = np.zeros(len(rows))
    lib().orc_accel_rows(len(rows), _p(rows), n, _p(pos), _p(h), n, _p(pos), _p(m), _p(h),
                         _p(acc), _p(phi))
    return acc, phi


def integrate(istep, dt, t, acc, pos_t, vel_t, pos, vel, acc0=None, ext=None):
    """kernelNbody_integTime; pos_t/vel_t/acc0 are updated in place."""
    n = pos.shape[1]
    lib().orc_integrate(n, istep, dt, t, ctypes.byref(ext) if ext is not None else None,
                        _p(acc), _p(pos_t), _p(vel_t), _p(pos), _p(vel), _p(acc0))


def diag(pos, vel, mass, phi=None):
    pos, vel, mass = _f64(pos), _f64(vel), _f64(mass)
    phi = None if phi is None else _f64(phi)
    out = np.zeros(17)
    lib().orc_diag(pos.shape[1], _p(pos), _p(vel), _p(mass), _p(phi), _p(out))
    return out


def run(pos, vel, mass, h, dt, steps, ext=None, compat_accumulate=False, with_diag=True,
        diag_every=1):
    """Reference step loop; returns (pos, vel, diag[(steps // diag_every + 1), 17]),
    diagnostics at t = 0 and after every diag_every-th step."""
    pos, vel = _f64(pos).copy(), _f64(vel).copy()
    mass, h = _f64(mass), _f64(h)
    d = np.zeros((steps // diag_every + 1, 17)) if with_diag else None
    lib().orc_run(pos.shape[1], _p(pos), _p(vel), _p(mass), _p(h), dt, steps,
                  ctypes.byref(ext) if ext is not None else None, 1 if compat_accumulate else 0,
                  _p(d), diag_every)
    return pos, vel, d


def p2mesh_bin(gmin, gcell, gsize, pos, mass):
    """Returns dict of cell_id, id_in_cell, count, offset, start_id_ref,
    block_incl, cell_mass."""
    gmin = _f64(gmin)
    gcell = _f64(gcell)
    gs = np.ascontiguousarray(gsize, dtype=np.int32)
    pos, mass = _f64(pos), _f64(mass)
    n = pos.shape[1]
    G = int(gs[0]) * int(gs[1]) * int(gs[2])
    B = (G + 511) // 512
    out = dict(cell_id=np.zeros(n, np.int32), id_in_cell=np.zeros(n, np.int32),
               count=np.zeros(G, np.int32), offset=np.zeros(G, np.int32),
               start_id_ref=np.zeros(G, np.int32), block_incl=np.zeros(B, np.int32),
               cell_mass=np.zeros(G))
    rc = lib().orc_p2mesh_bin(_p(gmin), _p(gcell), _p(gs), n, _p(pos), _p(mass),
                              _p(out["cell_id"]), _p(out["id_in_cell"]), _p(out["count"]),
                              _p(out["offset"]), _p(out["start_id_ref"]),
                              _p(out["block_incl"]), _p(out["cell_mass"]))
    if rc:
        raise ValueError("invalid grid")
    return out


def p2mesh_cell_phi(cell_id, count, phi, cell_phi=None, accumulate=False):
    """computeCellPhiUnsorted in index order; returns the new cell_phi."""
    cell_id = np.ascontiguousarray(cell_id, dtype=np.int32)
    count = np.ascontiguousarray(count, dtype=np.int32)
    phi = _f64(phi)
    G = count.shape[0]
    out = np.zeros(G) if cell_phi is None else _f64(cell_phi).copy()
    lib().orc_p2mesh_cell_phi(cell_id.shape[0], G, _p(cell_id), _p(count), _p(phi), _p(out),
                              1 if accumulate else 0)
    return out


def p2mesh_cell_weighted(cell_id, cell_mass, mass, value):
    """apply_particle_to_cell_weighted in index order; returns cell_value."""
    cell_id = np.ascontiguousarray(cell_id, dtype=np.int32)
    cell_mass, mass, value = _f64(cell_mass), _f64(mass), _f64(value)
    out = np.zeros(cell_mass.shape[0])
    lib().orc_p2mesh_cell_weighted(cell_id.shape[0], cell_mass.shape[0], _p(cell_id),
                                   _p(cell_mass), _p(mass), _p(value), _p(out))
    return out


def p2mesh_cell_avg(cell_id, count, cell_value):
    """apply_cell_to_particles_avg: per-body cell_value / count (-1 if empty)."""
    cell_id = np.ascontiguousarray(cell_id, dtype=np.int32)
    count = np.ascontiguousarray(count, dtype=np.int32)
    cell_value = _f64(cell_value)
    out = np.zeros(cell_id.shape[0])
    lib().orc_p2mesh_cell_avg(cell_id.shape[0], _p(cell_id), _p(count), _p(cell_value), _p(out))
    return out


def acc_abs(v):
    """acc_abs of vec3 planes (3, n): |v_i| in FP64."""
    v = _f64(v)
    out = np.zeros(v.shape[1])
    lib().orc_acc_abs(v.shape[1], _p(v), _p(out))
    return out


def sat_trunc(v):
    return lib().orc_sat_trunc(float(v))


def wave_step(w, phi, phi_old, mass):
    """w: dict(nx, dx, c, dt, diss_base, diss_extra, sim_l, bc_l)."""
    phi, phi_old, mass = _f64(phi), _f64(phi_old), _f64(mass)
    out = np.zeros_like(phi)
    lib().orc_wave_step(w["nx"], w["dx"], w["c"], w["dt"], w["diss_base"], w["diss_extra"],
                        w["sim_l"], w["bc_l"], _p(out), _p(phi), _p(phi_old), _p(mass))
    return out


def wave_iterate(w, prev, curr, mass, iters):
    prev, curr, mass = _f64(prev), _f64(curr), _f64(mass)
    out = np.zeros_like(curr)
    prev_out = np.zeros_like(curr)
    lib().orc_wave_iterate(w["nx"], w["dx"], w["c"], w["dt"], w["diss_base"], w["diss_extra"],
                           w["sim_l"], w["bc_l"], _p(prev), _p(curr), _p(mass), iters, _p(out),
                           _p(prev_out))
    return out, prev_out


def reference_wave(nx, dx):
    """nbody.cu:1134-1143 constants for an nx^3 grid of cell dx."""
    c = 4574.337022617616
    sim_l = 0.5 * nx * dx
    return dict(nx=nx, dx=dx, c=c, dt=0.5 * dx / (c * np.sqrt(3)), diss_base=0.1,
                diss_extra=0.01, sim_l=sim_l, bc_l=sim_l / 10.0)


def num_threads():
    return lib().orc_num_threads()


def mesh_acc(nx, dx, phi, want_abs=True):
    """acc_field on the full 3-D grid: returns (acc (3, nx^3), |acc| or None)."""
    phi = _f64(phi)
    G = nx ** 3
    acc = np.zeros((3, G))
    ab = np.zeros(G) if want_abs else None
    lib().orc_mesh_acc(nx, float(dx), _p(phi), _p(acc), _p(ab))
    return acc, ab


def mesh_to_particles(cell_id, cell_acc, cell_phi):
    """Per-particle cell values: returns (acc (3, n), phi (n,))."""
    cell_id = np.ascontiguousarray(cell_id, dtype=np.int32)
    cell_acc = _f64(cell_acc)
    cell_phi = _f64(cell_phi)
    n = cell_id.shape[0]
    acc = np.zeros((3, n))
    phi = np.zeros(n)
    lib().orc_mesh_to_particles(n, cell_phi.shape[0], _p(cell_id), _p(cell_acc), _p(cell_phi),
                                _p(acc), _p(phi))
    return acc, phi
