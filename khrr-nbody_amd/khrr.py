"""Python binding of the gfx950 N-body hot path (include/khrr.h) over ctypes.

The operators mirror the reference's kernels one to one:

  ==========================  ============================================
  Context.accel / .phi        ACCEL_kernel / PHI_kernel
                              (nbody/include/nbody-kernel.cuh:89-142, 43-86)
  Context.integrate           kernelNbody_integTime (nbody-kernel.cuh:144-194)
  extpot_init                 host constants nbody.cu:1111-1120 (dev_fex)
  Context.diag                result() sums, nbody.cu:797-833
  Context.p2mesh_bin          convert_particles_to_grid, nbody.cu:139-181
  Context.wave_step/_iterate  wave_diss_iteration + swap, nbody.cu:227-239
  ==========================  ============================================

Device buffers are torch tensors (torch supplies device memory and the
stream; all arithmetic runs in libkhrr.so).  "vec3 planes" are tensors of
shape (3, n): row 0 = x, row 1 = y, row 2 = z.  Every call raises
KhrrError with the library's message when the C-ABI returns an error, and
importing the module raises when the built library is missing: there is
no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import re
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libkhrr.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "khrr.h")

ACCUMULATE = 1
COMPENSATED = 2
GRAPH = 4

K_GRAVITY, K_INTEGRATE, K_BIN, K_WAVE, K_DIAG = range(5)
DIAG_NAMES = ("Ek", "Ep", "Px", "Py", "Pz", "Lx", "Ly", "Lz", "M", "Mx", "My", "Mz",
              "Vr_max", "Vfi_max", "Vz_max", "R_max", "Z_max")
BIN_BLOCK = 512


class KhrrError(RuntimeError):
    """Error status returned by the C-ABI."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"khrr error {code}: {msg}")
        self.code = code


class Tune(ctypes.Structure):
    _fields_ = [("i_per_lane", ctypes.c_int), ("waves_per_block", ctypes.c_int),
                ("j_split", ctypes.c_int), ("j_tile", ctypes.c_int), ("symmetric", ctypes.c_int),
                ("superblock", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ExtPot(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("Mh", "Mh_inf", "a", "Rh", "Mb", "b", "Rb", "Rh2", "con", "const1",
                 "c_phi_h", "c_phi_b")]


class Grid(ctypes.Structure):
    _fields_ = [("min", ctypes.c_double * 3), ("cell", ctypes.c_double * 3),
                ("size", ctypes.c_int * 3)]


class Wave(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("dx", ctypes.c_double), ("c", ctypes.c_double),
                ("dt", ctypes.c_double), ("diss_base", ctypes.c_double),
                ("diss_extra", ctypes.c_double), ("sim_l", ctypes.c_double),
                ("bc_l", ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("pos", ctypes.c_void_p), ("vel", ctypes.c_void_p),
                ("mass", ctypes.c_void_p), ("eps", ctypes.c_void_p), ("dt", ctypes.c_double),
                ("steps", ctypes.c_int), ("grid", Grid)]


class ServerStats(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int), ("pending", ctypes.c_int), ("finished", ctypes.c_int),
                ("rounds", ctypes.c_longlong), ("pair_interactions", ctypes.c_double),
                ("bytes_in_use", ctypes.c_size_t), ("bytes_budget", ctypes.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SIGS = {
    "khrr_abi_version": (_I, []),
    "khrr_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "khrr_ctx_destroy": (None, [_P]),
    "khrr_ctx_set_stream": (_I, [_P, _P]),
    "khrr_last_error": (ctypes.c_char_p, [_P]),
    "khrr_ctx_sync": (_I, [_P]),
    "khrr_ctx_set_timing": (_I, [_P, _I]),
    "khrr_ctx_kernel_time": (_I, [_P, _I, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(_D)]),
    "khrr_ctx_reset_timing": (_I, [_P]),
    "khrr_accel": (_I, [_P, _I, _P, _P, _I, _P, _P, _P, _P, _P, ctypes.c_uint,
                        ctypes.POINTER(Tune)]),
    "khrr_accel_batched": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, ctypes.c_uint,
                                ctypes.POINTER(Tune)]),
    "khrr_tune_gravity": (_I, [_P, _I, ctypes.POINTER(Tune), ctypes.POINTER(_D)]),
    "khrr_tune_default": (_I, [_P, _I, ctypes.POINTER(Tune)]),
    "khrr_tune_batched": (_I, [_P, _I, _P, ctypes.POINTER(Tune), ctypes.POINTER(_D)]),
    "khrr_tune_cached": (_I, [_P, _I, ctypes.POINTER(Tune)]),
    "khrr_tune_batched_cached": (_I, [_P, ctypes.POINTER(Tune)]),
    "khrr_tune_save": (_I, [_P, ctypes.c_char_p]),
    "khrr_tune_load": (_I, [_P, ctypes.c_char_p]),
    "khrr_extpot_init": (_I, [_D, _D, _D, _D, _D, _D, ctypes.POINTER(ExtPot)]),
    "khrr_extpot_force": (_I, [ctypes.POINTER(ExtPot), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "khrr_integrate": (_I, [_P, _I, _I, _D, _D, ctypes.POINTER(ExtPot), _P, _P, _P, _P, _P,
                            _P, _P]),
    "khrr_diag": (_I, [_P, _I, _P, _P, _P, _P, ctypes.POINTER(_D)]),
    "khrr_p2mesh_bin": (_I, [_P, ctypes.POINTER(Grid), _I, _P, _P, _P, _P, _P, _P, _P, _P,
                             _P]),
    "khrr_p2mesh_bin_noalloc": (_I, [_P, ctypes.POINTER(Grid), _I, _P, _P, _P, _P, _P, _P, _P,
                                     _P, _P]),
    "khrr_p2mesh_cell_phi": (_I, [_P, _I, ctypes.c_longlong, _P, _P, _P, _P, _P, _P, _I]),
    "khrr_mesh_acc": (_I, [_P, _I, _D, _P, _P, _P]),
    "khrr_mesh_to_particles": (_I, [_P, _I, ctypes.c_longlong, _P, _P, _P, _P, _P]),
    "khrr_p2mesh_cell_weighted": (_I, [_P, _I, ctypes.c_longlong, _P, _P, _P, _P, _P, _P, _P,
                                       _P]),
    "khrr_p2mesh_cell_avg": (_I, [_P, _I, ctypes.c_longlong, _P, _P, _P, _P]),
    "khrr_mesh_acc_f64": (_I, [_P, _I, _D, _P, _P, _P]),
    "khrr_acc_abs": (_I, [_P, _I, _P, _P]),
    "khrr_server_create": (_I, [_P, ctypes.c_size_t, _I, ctypes.POINTER(_P)]),
    "khrr_server_destroy": (None, [_P]),
    "khrr_server_set_round_pairs": (_I, [_P, _D]),
    "khrr_server_submit": (_I, [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_longlong)]),
    "khrr_server_step": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), _I, ctypes.POINTER(_I)]),
    "khrr_server_result": (_I, [_P, ctypes.c_longlong, ctypes.POINTER(_D), ctypes.POINTER(_D),
                                ctypes.POINTER(_D), ctypes.c_size_t]),
    "khrr_server_stats": (_I, [_P, ctypes.POINTER(ServerStats)]),
    "khrr_wave_source": (_I, [_P, ctypes.POINTER(Wave), _P, _P]),
    "khrr_wave_step": (_I, [_P, ctypes.POINTER(Wave), _P, _P, _P, _P]),
    "khrr_wave_iterate": (_I, [_P, ctypes.POINTER(Wave), ctypes.POINTER(_P), ctypes.POINTER(_P),
                               ctypes.POINTER(_P), _P, _I]),
}


def header_functions(path: str = HEADER_PATH) -> list[str]:
    """Names of the entry points include/khrr.h declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(khrr_[a-z0-9_]+)\s*\(", text)))


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"libkhrr.so not built at {path}; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _ptr(t, dtype=None, shape=None, name="tensor"):
    """Device pointer of a contiguous torch tensor (None -> NULL)."""
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a device tensor, got {t.device}")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    return t.data_ptr()


def extpot_init(Mh, a, Rh, Mb, b, Rb) -> ExtPot:
    e = ExtPot()
    rc = lib().khrr_extpot_init(Mh, a, Rh, Mb, b, Rb, ctypes.byref(e))
    if rc:
        raise KhrrError(rc, lib().khrr_last_error(None).decode())
    return e


def extpot_force(ext: ExtPot, p):
    pin = (_D * 3)(*p)
    out = (_D * 3)()
    rc = lib().khrr_extpot_force(ctypes.byref(ext), pin, out)
    if rc:
        raise KhrrError(rc, lib().khrr_last_error(None).decode())
    return list(out)


class Context:
    """One HIP device + stream + scratch arena (khrr_ctx)."""

    def __init__(self, device: int = 0, stream=None):
        self._lib = lib()
        self._servers = weakref.WeakSet()
        h = _P()
        rc = self._lib.khrr_ctx_create(device, ctypes.byref(h))
        if rc:
            raise KhrrError(rc, self._lib.khrr_last_error(None).decode())
        self._h = h
        self.device = device
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        # servers hold the context: they go first
        for srv in list(getattr(self, "_servers", ())):
            srv.close()
        if getattr(self, "_h", None):
            self._lib.khrr_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc:
            raise KhrrError(rc, self._lib.khrr_last_error(self._h).decode())

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream, a raw hipStream_t int, or None."""
        raw = getattr(stream, "cuda_stream", stream)
        self._check(self._lib.khrr_ctx_set_stream(self._h, raw))

    def sync(self):
        self._check(self._lib.khrr_ctx_sync(self._h))

    # -- timing --------------------------------------------------------
    def set_timing(self, on: bool):
        self._check(self._lib.khrr_ctx_set_timing(self._h, 1 if on else 0))

    def reset_timing(self):
        self._check(self._lib.khrr_ctx_reset_timing(self._h))

    def kernel_time(self, cls: int):
        n = ctypes.c_long()
        ms = _D()
        self._check(self._lib.khrr_ctx_kernel_time(self._h, cls, ctypes.byref(n),
                                                   ctypes.byref(ms)))
        return n.value, ms.value

    # -- gravity -------------------------------------------------------
    def accel(self, pos_i, h_i, pos_j, m_j, h_j, acc=None, phi=None, accumulate=False,
              compensated=False, tune: Tune | None = None):
        """ACCEL_kernel / PHI_kernel.  pos_*: float32 (3, n); h = eps^2/2."""
        import torch
        f32 = torch.float32
        n_i = pos_i.shape[1]
        n_j = pos_j.shape[1]
        flags = (ACCUMULATE if accumulate else 0) | (COMPENSATED if compensated else 0)
        self._check(self._lib.khrr_accel(
            self._h, n_i, _ptr(pos_i, f32, (3, n_i), "pos_i"), _ptr(h_i, f32, (n_i,), "h_i"),
            n_j, _ptr(pos_j, f32, (3, n_j), "pos_j"), _ptr(m_j, f32, (n_j,), "m_j"),
            _ptr(h_j, f32, (n_j,), "h_j"), _ptr(acc, f32, (3, n_i), "acc"),
            _ptr(phi, f32, (n_i,), "phi"), flags,
            ctypes.byref(tune) if tune is not None else None))

    def accel_batched(self, sys_off_dev, sys_off_host, pos, h, m, acc=None, phi=None,
                      accumulate=False, compensated=False, tune: Tune | None = None,
                      graph=False):
        """graph=True: capture/replay the launch sequence as a HIP graph."""
        import torch
        f32 = torch.float32
        n_sys = len(sys_off_host) - 1
        n = int(sys_off_host[-1])
        host = (ctypes.c_int * (n_sys + 1))(*[int(v) for v in sys_off_host])
        flags = ((ACCUMULATE if accumulate else 0) | (COMPENSATED if compensated else 0) |
                 (GRAPH if graph else 0))
        self._check(self._lib.khrr_accel_batched(
            self._h, n_sys, _ptr(sys_off_dev, torch.int32, (n_sys + 1,), "sys_off"),
            ctypes.cast(host, _P), _ptr(pos, f32, (3, n), "pos"), _ptr(h, f32, (n,), "h"),
            _ptr(m, f32, (n,), "m"), _ptr(acc, f32, (3, n), "acc"), _ptr(phi, f32, (n,), "phi"),
            flags, ctypes.byref(tune) if tune is not None else None))

    def tune_gravity(self, n: int):
        t = Tune()
        ms = _D()
        self._check(self._lib.khrr_tune_gravity(self._h, n, ctypes.byref(t), ctypes.byref(ms)))
        return t, ms.value

    def tune_batched(self, sys_off_host):
        n_sys = len(sys_off_host) - 1
        host = (ctypes.c_int * (n_sys + 1))(*[int(v) for v in sys_off_host])
        t = Tune()
        ms = _D()
        self._check(self._lib.khrr_tune_batched(self._h, n_sys, ctypes.cast(host, _P),
                                                ctypes.byref(t), ctypes.byref(ms)))
        return t, ms.value

    def tune_default(self, n: int) -> Tune:
        t = Tune()
        self._check(self._lib.khrr_tune_default(self._h, n, ctypes.byref(t)))
        return t

    def tune_cached(self, n: int):
        """The tuned shape khrr_accel uses for n bodies, or None."""
        t = Tune()
        rc = self._lib.khrr_tune_cached(self._h, n, ctypes.byref(t))
        if rc == -4:  # KHRR_ESTATE: nothing tuned for this size bucket
            return None
        self._check(rc)
        return t

    def tune_batched_cached(self):
        t = Tune()
        rc = self._lib.khrr_tune_batched_cached(self._h, ctypes.byref(t))
        if rc == -4:
            return None
        self._check(rc)
        return t

    def tune_save(self, path: str):
        """Writes the context's tuned shapes to a JSON cache file."""
        self._check(self._lib.khrr_tune_save(self._h, os.fsencode(path)))

    def tune_load(self, path: str):
        """Loads a JSON cache file made on the same kind of device."""
        self._check(self._lib.khrr_tune_load(self._h, os.fsencode(path)))

    # -- integrator ----------------------------------------------------
    def integrate(self, istep, dt, t, acc, pos_t, vel_t, pos, vel, acc0=None, pos32=None,
                  ext: ExtPot | None = None):
        import torch
        n = pos.shape[1]
        f32, f64 = torch.float32, torch.float64
        self._check(self._lib.khrr_integrate(
            self._h, n, istep, dt, t, ctypes.byref(ext) if ext is not None else None,
            _ptr(acc, f32, (3, n), "acc"), _ptr(pos_t, f64, (3, n), "pos_t"),
            _ptr(vel_t, f64, (3, n), "vel_t"), _ptr(pos, f64, (3, n), "pos"),
            _ptr(vel, f64, (3, n), "vel"), _ptr(acc0, f32, (3, n), "acc0"),
            _ptr(pos32, f32, (3, n), "pos32")))

    # -- diagnostics ---------------------------------------------------
    def diag(self, pos, vel, mass, phi=None) -> dict:
        import torch
        n = pos.shape[1]
        out = (_D * len(DIAG_NAMES))()
        self._check(self._lib.khrr_diag(
            self._h, n, _ptr(pos, torch.float64, (3, n), "pos"),
            _ptr(vel, torch.float64, (3, n), "vel"), _ptr(mass, torch.float64, (n,), "mass"),
            _ptr(phi, torch.float32, (n,), "phi"), out))
        return dict(zip(DIAG_NAMES, list(out)))

    # -- binning -------------------------------------------------------
    def p2mesh_bin(self, grid: Grid, pos, mass, cell_id=None, id_in_cell=None, count=None,
                   offset=None, start_id_ref=None, block_incl=None, cell_mass=None):
        import torch
        n = pos.shape[1]
        G = grid.size[0] * grid.size[1] * grid.size[2]
        B = (G + BIN_BLOCK - 1) // BIN_BLOCK
        i32 = torch.int32
        self._check(self._lib.khrr_p2mesh_bin(
            self._h, ctypes.byref(grid), n, _ptr(pos, torch.float64, (3, n), "pos"),
            _ptr(mass, torch.float64, (n,), "mass"), _ptr(cell_id, i32, (n,), "cell_id"),
            _ptr(id_in_cell, i32, (n,), "id_in_cell"), _ptr(count, i32, (G,), "count"),
            _ptr(offset, i32, (G,), "offset"), _ptr(start_id_ref, i32, (G,), "start_id_ref"),
            _ptr(block_incl, i32, (B,), "block_incl"),
            _ptr(cell_mass, torch.float64, (G,), "cell_mass")))

    def p2mesh_bin_noalloc(self, grid: Grid, pos, mass, cell_id=None, id_in_cell=None,
                           count=None, offset=None, start_id_ref=None, block_incl=None,
                           cell_mass=None):
        """p2mesh_bin with the library's fixed scratch (allocates nothing;
        at most 1048576 bodies on 256^3 cells)."""
        import torch
        n = pos.shape[1]
        G = grid.size[0] * grid.size[1] * grid.size[2]
        B = (G + BIN_BLOCK - 1) // BIN_BLOCK
        i32 = torch.int32
        self._check(self._lib.khrr_p2mesh_bin_noalloc(
            self._h, ctypes.byref(grid), n, _ptr(pos, torch.float64, (3, n), "pos"),
            _ptr(mass, torch.float64, (n,), "mass"), _ptr(cell_id, i32, (n,), "cell_id"),
            _ptr(id_in_cell, i32, (n,), "id_in_cell"), _ptr(count, i32, (G,), "count"),
            _ptr(offset, i32, (G,), "offset"), _ptr(start_id_ref, i32, (G,), "start_id_ref"),
            _ptr(block_incl, i32, (B,), "block_incl"),
            _ptr(cell_mass, torch.float64, (G,), "cell_mass")))

    def p2mesh_cell_phi(self, cell_id, id_in_cell, count, offset, phi, cell_phi,
                        accumulate=False):
        """computeCellPhiUnsorted from the outputs of p2mesh_bin."""
        import torch
        n = cell_id.shape[0]
        G = count.shape[0]
        i32 = torch.int32
        self._check(self._lib.khrr_p2mesh_cell_phi(
            self._h, n, G, _ptr(cell_id, i32, (n,), "cell_id"),
            _ptr(id_in_cell, i32, (n,), "id_in_cell"), _ptr(count, i32, (G,), "count"),
            _ptr(offset, i32, (G,), "offset"), _ptr(phi, torch.float32, (n,), "phi"),
            _ptr(cell_phi, torch.float64, (G,), "cell_phi"), 1 if accumulate else 0))

    def p2mesh_cell_weighted(self, cell_id, id_in_cell, count, offset, cell_mass, mass, value,
                             cell_value):
        """apply_particle_to_cell_weighted in index order: cell_value[c] =
        sum over the cell's bodies of value[i] * (mass[i] / cell_mass[c])."""
        import torch
        n = cell_id.shape[0]
        G = count.shape[0]
        i32, f64 = torch.int32, torch.float64
        self._check(self._lib.khrr_p2mesh_cell_weighted(
            self._h, n, G, _ptr(cell_id, i32, (n,), "cell_id"),
            _ptr(id_in_cell, i32, (n,), "id_in_cell"), _ptr(count, i32, (G,), "count"),
            _ptr(offset, i32, (G,), "offset"), _ptr(cell_mass, f64, (G,), "cell_mass"),
            _ptr(mass, f64, (n,), "mass"), _ptr(value, f64, (n,), "value"),
            _ptr(cell_value, f64, (G,), "cell_value")))

    def p2mesh_cell_avg(self, cell_id, count, cell_value, out):
        """apply_cell_to_particles_avg: out[i] = cell_value[c] / count[c]
        (-1 for an empty cell, NaN for an index outside the grid)."""
        import torch
        n = cell_id.shape[0]
        G = count.shape[0]
        self._check(self._lib.khrr_p2mesh_cell_avg(
            self._h, n, G, _ptr(cell_id, torch.int32, (n,), "cell_id"),
            _ptr(count, torch.int32, (G,), "count"),
            _ptr(cell_value, torch.float64, (G,), "cell_value"),
            _ptr(out, torch.float64, (n,), "out")))

    # -- stencil -------------------------------------------------------
    def wave_source(self, wave: Wave, cell_mass, src):
        import torch
        G = wave.nx ** 3
        self._check(self._lib.khrr_wave_source(
            self._h, ctypes.byref(wave), _ptr(cell_mass, torch.float64, (G,), "cell_mass"),
            _ptr(src, torch.float32, (G,), "src")))

    def wave_step(self, wave: Wave, phi_new, phi, phi_old, src):
        import torch
        G = wave.nx ** 3
        f32 = torch.float32
        self._check(self._lib.khrr_wave_step(
            self._h, ctypes.byref(wave), _ptr(phi_new, f32, (G,), "phi_new"),
            _ptr(phi, f32, (G,), "phi"), _ptr(phi_old, f32, (G,), "phi_old"),
            _ptr(src, f32, (G,), "src")))

    def wave_iterate(self, wave: Wave, prev, curr, nxt, src, iters: int):
        """Runs iters steps with the reference rotation; returns the tensors
        (prev, curr, next) in their final roles (curr = newest field)."""
        import torch
        G = wave.nx ** 3
        f32 = torch.float32
        bufs = {t.data_ptr(): t for t in (prev, curr, nxt)}
        p = _P(_ptr(prev, f32, (G,), "prev"))
        c = _P(_ptr(curr, f32, (G,), "curr"))
        x = _P(_ptr(nxt, f32, (G,), "next"))
        self._check(self._lib.khrr_wave_iterate(
            self._h, ctypes.byref(wave), ctypes.byref(p), ctypes.byref(c), ctypes.byref(x),
            _ptr(src, f32, (G,), "src"), iters))
        return bufs[p.value], bufs[c.value], bufs[x.value]

    # -- mesh force of the wave solver ------------------------------------
    def mesh_acc(self, nx: int, dx: float, cell_phi, cell_acc, cell_acc_abs=None):
        """acc_field: cell_acc (3, nx^3) float32 = -grad(cell_phi) (nx^3 float32)."""
        import torch
        f32 = torch.float32
        G = nx ** 3
        self._check(self._lib.khrr_mesh_acc(
            self._h, int(nx), float(dx), _ptr(cell_phi, f32, (G,), "cell_phi"),
            _ptr(cell_acc, f32, (3, G), "cell_acc"),
            _ptr(cell_acc_abs, f32, (G,), "cell_acc_abs")))

    def mesh_acc_f64(self, nx: int, dx: float, cell_phi, cell_acc, cell_acc_abs=None):
        """acc_field in FP64: cell_acc (3, nx^3) float64 of cell_phi (nx^3 float64)."""
        import torch
        f64 = torch.float64
        G = nx ** 3
        self._check(self._lib.khrr_mesh_acc_f64(
            self._h, int(nx), float(dx), _ptr(cell_phi, f64, (G,), "cell_phi"),
            _ptr(cell_acc, f64, (3, G), "cell_acc"),
            _ptr(cell_acc_abs, f64, (G,), "cell_acc_abs")))

    def acc_abs(self, v, out):
        """acc_abs: out[i] = |v[:, i]| in FP64 of float32 vec3 planes (3, n)."""
        import torch
        n = v.shape[-1]
        self._check(self._lib.khrr_acc_abs(
            self._h, n, _ptr(v, torch.float32, (3, n), "v"),
            _ptr(out, torch.float64, (n,), "out")))

    def mesh_to_particles(self, cell_id, cell_acc=None, cell_phi=None, acc=None, phi=None):
        """acc[:, i] = cell_acc[:, cell_id[i]], phi[i] = cell_phi[cell_id[i]]."""
        import torch
        n = cell_id.shape[0]
        G = (cell_phi if cell_phi is not None else cell_acc).shape[-1]
        f32 = torch.float32
        self._check(self._lib.khrr_mesh_to_particles(
            self._h, n, G, _ptr(cell_id, torch.int32, (n,), "cell_id"),
            _ptr(cell_acc, f32, (3, G), "cell_acc"), _ptr(cell_phi, f32, (G,), "cell_phi"),
            _ptr(acc, f32, (3, n), "acc"), _ptr(phi, f32, (n,), "phi")))


def reference_wave(nx: int, dx: float) -> Wave:
    """The reference's wave constants for a grid of nx cells of size dx
    (nbody.cu:1134-1143): c = 4574.337022617616, dt = dx/(2 c sqrt 3),
    diss_base 0.1, diss_extra 0.01, sim_l = nx*dx/2, bc_l = sim_l/10."""
    import math
    c = 4574.337022617616
    sim_l = 0.5 * nx * dx
    return Wave(nx, dx, c, 0.5 * dx / (c * math.sqrt(3)), 0.1, 0.01, sim_l, sim_l / 10.0)


def cube_grid(nx: int, lo: float, hi: float) -> Grid:
    """Cubic grid of nx^3 cells over [lo, hi]^3 (WaveEqData, nbody.cu:1146-1165)."""
    d = (hi - lo) / nx
    return Grid((ctypes.c_double * 3)(lo, lo, lo), (ctypes.c_double * 3)(d, d, d),
                (ctypes.c_int * 3)(nx, nx, nx))


class Server:
    """Continuous-batching server (khrr_server_*): admits simulations by
    device-memory footprint, keeps their state on the device and advances
    all of them with one batched gravity launch per step."""

    def __init__(self, ctx: Context, budget_bytes: int, max_active: int,
                 round_pairs: float = 0.0):
        self._ctx = ctx
        self._lib = ctx._lib
        h = _P()
        ctx._check(self._lib.khrr_server_create(ctx._h, int(budget_bytes), int(max_active),
                                                ctypes.byref(h)))
        self._h = h
        ctx._servers.add(self)
        if round_pairs:
            self.set_round_pairs(round_pairs)

    def set_round_pairs(self, pairs: float):
        """0: every active simulation steps each round; > 0: least remaining
        work first until a round holds `pairs` pair-interactions."""
        self._ctx._check(self._lib.khrr_server_set_round_pairs(self._h, float(pairs)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.khrr_server_destroy(self._h)
            self._h = None
            self._ctx._servers.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, pos, vel, mass, eps, dt: float, steps: int, grid: Grid) -> int:
        """pos, vel: float64 (3, n); mass, eps: float64 (n,) host arrays."""
        import numpy as np
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        n = pos.shape[1]
        if pos.shape != (3, n) or vel.shape != (3, n) or mass.shape != (n,) or eps.shape != (n,):
            raise ValueError("request arrays must be (3, n), (3, n), (n,), (n,)")
        req = Request(n, pos.ctypes.data, vel.ctypes.data, mass.ctypes.data, eps.ctypes.data,
                      float(dt), int(steps), grid)
        rid = ctypes.c_longlong()
        self._ctx._check(self._lib.khrr_server_submit(self._h, ctypes.byref(req),
                                                      ctypes.byref(rid)))
        return rid.value

    def step(self, max_done: int = 4096):
        """One scheduling round; returns the ids that finished in it.  A
        round that admits or finishes a simulation returns when the device
        has completed it; others return once queued (at most two in
        flight), so time the server from the step that finds work to the
        step that drains it."""
        ids = (ctypes.c_longlong * max_done)()
        nd = ctypes.c_int()
        self._ctx._check(self._lib.khrr_server_step(self._h, ids, max_done, ctypes.byref(nd)))
        return [ids[k] for k in range(min(nd.value, max_done))]

    def result(self, rid: int, cells: int):
        import numpy as np
        d0 = (_D * len(DIAG_NAMES))()
        d1 = (_D * len(DIAG_NAMES))()
        cm = np.zeros(cells, dtype=np.float64)
        self._ctx._check(self._lib.khrr_server_result(
            self._h, rid, d0, d1, cm.ctypes.data_as(ctypes.POINTER(_D)), cells))
        return dict(zip(DIAG_NAMES, list(d0))), dict(zip(DIAG_NAMES, list(d1))), cm

    def stats(self) -> dict:
        st = ServerStats()
        self._ctx._check(self._lib.khrr_server_stats(self._h, ctypes.byref(st)))
        return st.as_dict()
