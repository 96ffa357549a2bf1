#!/usr/bin/env python3
"""Headline benchmark of the MI355X N-body hot path.

Workload (one GPU): one disk + halo galaxy, 262144 stars (mass 1.0) and
524288 halo bodies (mass 6.699356), eps 0.004, dt 0.002, FP32 direct-sum
gravity with FP64 predictor-corrector state.  A step is the reference's
predictor -> ACCEL -> corrector (nbody.cu:1525-1582) over all 786432 bodies,
50 timed steps by default after 2 warm-up steps; `value` is whole-step
pair-interactions per second (N^2 per step, i = j counted as the reference
does), inputs resident in HBM.

Also reported on the same line (one GPU):
  roofline      gravity kernel: 20 FLOP per pair vs the 157.3 TFLOP/s FP32
                peak, average launch time from HIP events on the library's
                stream; traffic from rocprof counters when supplied
  binning       p2mesh of the final state on 256^3 over [-2,2]^3
  stencil       200 damped-wave iterations on that grid from phi = 0
  mesh          the mesh-side entry points on that grid: the wave solver's
                mesh force, the bodies' cell potential, the debug transfers
  batched       512 Plummer spheres (N log-uniform 1024..131072) per step,
                one persistent launch vs one plain launch per system
  serving       a Poisson stream of mixed requests through the server
                (two round schedules) vs the same requests one at a time
  precision     3.1M-body two-galaxy encounter: plain vs compensated sums,
                time per evaluation and drift over a few steps
  drift         E, P, L relative drift over the timed steps (FP64 device
                reductions)
  cpu leg       last, unless --no-cpu: the FP64 oracle (OpenMP, all host
                cores) timed on 16384 rows of the headline workload
                (cpu_baseline), and as the checker of the GPU legs the FP64
                references behind `accuracy` (headline forces), the
                stencil's err_vs_fp64 and the precision form's errors

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "khrr-nbody_amd"))

METRIC = "G pair-interactions/s & % FP32 VALU peak; batched/served vs plain; p99 lat"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOP_PER_PAIR = 20
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
TUNE_CACHE = os.path.join(ROOT, "khrr-nbody_amd", "tune_cache.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)  # the headline config's 50 steps
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-stars", type=int, default=262144)
    ap.add_argument("--n-dm", type=int, default=524288)
    ap.add_argument("--cpu-rows", type=int, default=16384)
    ap.add_argument("--quick", action="store_true", help="headline only (no extras)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=2)
    ap.add_argument("--batched-run-steps", type=int, default=1000,
                    help="steps of the batched-run leg (512 systems advanced together)")
    ap.add_argument("--serve-requests", type=int, default=16,
                    help="requests of the serving leg (the full stream is 5000)")
    ap.add_argument("--serve-lambda", type=float, default=1.2, help="arrivals per second")
    ap.add_argument("--serve-tmax", default="0.2,2.0",
                    help="range of the requests' integration times tmax (uniform), lo,hi")
    ap.add_argument("--serve-saturated", type=float, default=50.0,
                    help="arrival rate of the saturated serving run (0: skip it)")
    ap.add_argument("--serve-round-pairs", type=float, default=3.4e10,
                    help="pair-interactions per step of a least-remaining-work round "
                         "(khrr_server_set_round_pairs)")
    ap.add_argument("--legs", default="binning,stencil,mesh,batched,serving,precision",
                    help="extras to run after the headline (comma list; also batched-run, "
                         "the 512 systems advanced --batched-run-steps steps)")
    ap.add_argument("--precision-steps", type=int, default=3,
                    help="steps of the 3.1M-body encounter per sum mode (0: skip)")
    ap.add_argument("--shape", default=None,
                    help="fixed launch shape i_per_lane,waves_per_block,j_split,j_tile,symmetric"
                         "[,superblock] instead of the autotuner (for profiler passes)")
    ap.add_argument("--tune-cache", default=TUNE_CACHE,
                    help="autotuner cache (JSON) loaded before tuning; shapes it holds for "
                         "this device are used as they are ('' = none)")
    ap.add_argument("--retune", action="store_true",
                    help="time the candidate shapes again instead of using the cache")
    ap.add_argument("--tune-save", default=None,
                    help="write the context's tuned shapes to this JSON file at the end")
    ap.add_argument("--stream-requests", type=int, default=5000,
                    help="length of the serving stream (--stream-range runs a segment of it)")
    ap.add_argument("--stream-range", default=None,
                    help="first,last: serve requests [first, last) of the stream in real time "
                         "and write their records to --stream-out (no other leg runs)")
    ap.add_argument("--stream-out", default="stream_segment.json",
                    help="where --stream-range writes its per-request records")
    ap.add_argument("--stream-fifo", type=int, default=0,
                    help="also run the segment's first K requests one at a time (plain FIFO)")
    ap.add_argument("--stream-fifo-only", action="store_true",
                    help="run the --stream-range requests one at a time only (plain FIFO, no "
                         "served run); --stream-merge takes the record for the FIFO figures")
    ap.add_argument("--stream-merge", default=None,
                    help="comma list of --stream-out files of consecutive segments: print "
                         "their combined figures (no GPU)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 512 MiB device copy first (counter calibration under rocprof)")
    return ap.parse_args()


class Sim:
    """Device state of one system and the reference step sequence."""

    def __init__(self, ctx, pos, vel, mass, h, dt):
        import torch
        dev = torch.device("cuda:0")
        f64 = dict(device=dev, dtype=torch.float64)
        f32 = dict(device=dev, dtype=torch.float32)
        self.ctx, self.dt, self.n = ctx, dt, pos.shape[1]
        self.pos = torch.tensor(pos, **f64)
        self.vel = torch.tensor(vel, **f64)
        self.mass64 = torch.tensor(mass, **f64)
        self.m32 = torch.tensor(mass, **f32)
        self.h32 = torch.tensor(h, **f32)
        self.pos32 = self.pos.float().contiguous()
        n = self.n
        self.post = torch.zeros((3, n), **f64)
        self.velt = torch.zeros((3, n), **f64)
        self.acc = torch.zeros((3, n), **f32)
        self.acct = torch.zeros((3, n), **f32)
        self.phi = torch.zeros(n, **f32)
        self.t = 0.0
        self.tune = None  # launch shape; None = the context's tuned/static choice
        self.compensated = False

    def gravity(self, out, phi=None):
        self.ctx.accel(self.pos32, self.h32, self.pos32, self.m32, self.h32, acc=out, phi=phi,
                       tune=self.tune, compensated=self.compensated)

    def init_forces(self):
        self.gravity(self.acc)

    def step(self):
        c = self.ctx
        c.integrate(0, self.dt, self.t, self.acc, self.post, self.velt, self.pos, self.vel,
                    pos32=self.pos32)
        self.gravity(self.acct)
        c.integrate(1, self.dt, self.t + self.dt, self.acct, self.pos, self.vel, self.post,
                    self.velt, acc0=self.acc)
        self.t += self.dt

    def diagnostics(self):
        self.ctx.accel(self.pos32, self.h32, self.pos32, self.m32, self.h32, phi=self.phi,
                       tune=self.tune, compensated=self.compensated)
        return self.ctx.diag(self.pos, self.vel, self.mass64, self.phi)


def energy(d):
    return 0.5 * (d["Ek"] + d["Ep"])


def hbm_figure(ms, alg_bytes):
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    return {"ms": ms, "alg_bytes": alg_bytes, "alg_GBps": gbs, "frac_hbm_peak": gbs / PEAK_HBM_GBS}


def bench_precision(ctx, steps, check, samples=8192):
    """Two-galaxy encounter (2 x (524288 + 1048576) = 3145728 bodies): plain
    FP32 sums vs compensated (TwoSum) sums: time per force evaluation and
    FP64 energy, momentum and angular-momentum drift over `steps`
    predictor-corrector steps.  The forces of `samples` random bodies at
    t = 0 go to `check` for the CPU leg's FP64 comparison."""
    import numpy as np
    import torch
    import ics
    import khrr
    pos, vel, mass, h = ics.encounter()
    n = pos.shape[1]
    rows = np.sort(np.random.default_rng(samples).choice(n, samples, replace=False))
    check.update(rows=rows, pos=pos.astype(np.float32).astype(np.float64),
                 m=mass.astype(np.float32).astype(np.float64),
                 h=h.astype(np.float32).astype(np.float64), acc={})
    out = {"workload": "two-galaxy encounter, 2 x (524288 + 1048576) bodies, X = +-2, "
                       "tilts 0/45 deg, V = (-+0.3, +-0.05)",
           "n_bodies": n, "samples": samples, "steps": steps}
    for mode in ("plain", "compensated"):
        sim = Sim(ctx, pos, vel, mass, h, 0.002)
        sim.compensated = mode == "compensated"
        torch.cuda.synchronize()
        ctx.reset_timing()
        ctx.set_timing(True)
        sim.init_forces()
        torch.cuda.synchronize()
        launches, ms = ctx.kernel_time(khrr.K_GRAVITY)
        check["acc"][mode] = sim.acc.cpu().numpy()[:, rows].astype(np.float64)
        d0 = sim.diagnostics()
        for k in range(steps):
            sim.step()
            if steps >= 50 and (k + 1) % 25 == 0:
                torch.cuda.synchronize()
                log(f"precision {mode}: step {k + 1}/{steps}")
        d1 = sim.diagnostics()
        ctx.set_timing(False)
        torch.cuda.synchronize()
        out[mode] = {
            "force_ms": ms / max(launches, 1),
            "Gpairs_per_s": float(n) * n / (ms / max(launches, 1) * 1e-3) / 1e9,
            "dE_rel": energy(d1) / energy(d0) - 1.0,
            "dP_abs": math.sqrt(sum((d1[k] - d0[k]) ** 2 for k in ("Px", "Py", "Pz"))),
            "dLz_rel": d1["Lz"] / d0["Lz"] - 1.0}
        log(f"precision {mode}: {out[mode]['force_ms']:.1f} ms per force evaluation, "
            f"dE/E {out[mode]['dE_rel']:.2e} over {steps} steps")
        del sim
        torch.cuda.empty_cache()
    return out


def bench_binning(ctx, sim, reps=5):
    import torch
    import khrr
    nx = 256
    grid = khrr.cube_grid(nx, -2.0, 2.0)
    G = nx ** 3
    B = (G + 511) // 512
    n = sim.n
    i32 = dict(device="cuda:0", dtype=torch.int32)
    bufs = dict(cell_id=torch.zeros(n, **i32), id_in_cell=torch.zeros(n, **i32),
                count=torch.zeros(G, **i32), offset=torch.zeros(G, **i32),
                start_id_ref=torch.zeros(G, **i32), block_incl=torch.zeros(B, **i32),
                cell_mass=torch.zeros(G, device="cuda:0", dtype=torch.float64))
    ctx.p2mesh_bin(grid, sim.pos, sim.mass64, **bufs)
    torch.cuda.synchronize()
    ctx.reset_timing()
    for _ in range(reps):
        ctx.p2mesh_bin(grid, sim.pos, sim.mass64, **bufs)
    launches, ms = ctx.kernel_time(khrr.K_BIN)
    avg_ms = ms / max(launches, 1)
    # algorithmic bytes: per body pos (24) + mass (8) read, cell_id (4) +
    # id_in_cell (4) written; per cell count, offset, start_id_ref (4 each)
    # and cell_mass (8) written
    alg = 40 * n + 20 * G
    return bufs, {
        "grid": "256^3 over [-2,2]^3", "n": n, "ms": avg_ms, "alg_bytes": alg,
        "alg_GBps": alg / (avg_ms * 1e-3) / 1e9,
        "frac_hbm_peak": alg / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
        "occupied_cells": int((bufs["count"] > 0).sum().item())}


def bench_stencil(ctx, cell_mass, check, iters=200):
    """200 damped-wave iterations on 256^3 from phi = 0; the mass grid and
    the final field go to `check` for the CPU leg's FP64 comparison."""
    import numpy as np
    import torch
    import khrr
    nx = 256
    w = khrr.reference_wave(nx, 4.0 / nx)
    G = nx ** 3
    src = torch.zeros(G, device="cuda:0")
    ctx.wave_source(w, cell_mass, src)
    b = [torch.zeros(G, device="cuda:0") for _ in range(3)]
    ctx.wave_iterate(w, b[0], b[1], b[2], src, 5)
    for t in b:
        t.zero_()
    torch.cuda.synchronize()
    ctx.reset_timing()
    _, cur, _ = ctx.wave_iterate(w, b[0], b[1], b[2], src, iters)
    # one timed bracket per khrr_wave_iterate call: every launch and copy
    _, ms = ctx.kernel_time(khrr.K_WAVE)
    avg_ms = ms / iters
    alg = 16 * G
    check.update(nx=nx, iters=iters, cell_mass=cell_mass.cpu().numpy(),
                 phi=cur.cpu().numpy().astype(np.float64))
    return cur, {"grid": "256^3", "iterations": iters, "ms_per_iter": avg_ms,
                 "alg_bytes_per_iter": alg, "alg_GBps": alg / (avg_ms * 1e-3) / 1e9,
                 "frac_hbm_peak": alg / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                 "max_abs_phi": float(cur.abs().max().item())}


def bench_mesh(ctx, sim, bins, field, reps=10):
    """The mesh-side entry points on the binning and the stencil field of
    the headline grid (256^3, 786432 bodies): the wave solver's mesh force
    (acc_field and the transfer to the bodies), the bodies' cell potential,
    and the transfers of the debug dumps.  Algorithmic bytes are each entry
    point's own inputs and outputs (scratch excluded); body values that the
    legs before do not produce (the bodies' potential, a per-body value)
    are seeded random, which the timing does not depend on."""
    import numpy as np
    import torch
    import khrr
    nx = 256
    G = nx ** 3
    dx = 4.0 / nx
    n = sim.n
    rng = np.random.default_rng(7)
    f32 = dict(device="cuda:0", dtype=torch.float32)
    f64 = dict(device="cuda:0", dtype=torch.float64)
    phi_b = torch.as_tensor(rng.normal(size=n).astype(np.float32)).to("cuda:0")
    val_b = torch.as_tensor(rng.normal(size=n)).to("cuda:0")
    cacc = torch.zeros((3, G), **f32)
    cabs = torch.zeros(G, **f32)
    pacc = torch.zeros((3, n), **f32)
    pphi = torch.zeros(n, **f32)
    cphi64 = torch.zeros(G, **f64)
    cacc64 = torch.zeros((3, G), **f64)
    cabs64 = torch.zeros(G, **f64)
    cval = torch.zeros(G, **f64)
    bout = torch.zeros(n, **f64)
    b = bins
    legs = [
        ("mesh_acc", khrr.K_WAVE, 20 * G,
         lambda: ctx.mesh_acc(nx, dx, field, cacc, cabs)),
        ("mesh_to_particles", khrr.K_BIN, 36 * n,
         lambda: ctx.mesh_to_particles(b["cell_id"], cacc, field, pacc, pphi)),
        ("p2mesh_cell_phi", khrr.K_BIN, 12 * n + 16 * G,
         lambda: ctx.p2mesh_cell_phi(b["cell_id"], b["id_in_cell"], b["count"], b["offset"],
                                     phi_b, cphi64)),
        ("p2mesh_cell_weighted", khrr.K_BIN, 24 * n + 24 * G,
         lambda: ctx.p2mesh_cell_weighted(b["cell_id"], b["id_in_cell"], b["count"],
                                          b["offset"], b["cell_mass"], sim.mass64, val_b, cval)),
        ("p2mesh_cell_avg", khrr.K_BIN, 24 * n,
         lambda: ctx.p2mesh_cell_avg(b["cell_id"], b["count"], b["cell_mass"], bout)),
        ("acc_abs", khrr.K_BIN, 20 * n, lambda: ctx.acc_abs(pacc, bout)),
        ("mesh_acc_f64", khrr.K_WAVE, 40 * G,
         lambda: ctx.mesh_acc_f64(nx, dx, cphi64, cacc64, cabs64)),
    ]
    out = {}
    for name, cls, alg, call in legs:
        call()
        torch.cuda.synchronize()
        ctx.reset_timing()
        for _ in range(reps):
            call()
        launches, ms = ctx.kernel_time(cls)
        avg_ms = ms / max(launches, 1)
        gbs = alg / (avg_ms * 1e-3) / 1e9
        out[name] = {"ms": avg_ms, "alg_bytes": alg, "alg_GBps": gbs,
                     "frac_hbm_peak": gbs / PEAK_HBM_GBS}
    return {"grid": "256^3 over [-2,2]^3", "n": n, "entry_points": out}


def bench_batched(ctx, steps, retune=False):
    """512 Plummer spheres: one persistent batched launch per step vs one
    plain launch per system per step (same integrator calls)."""
    import numpy as np
    import torch
    import ics
    import khrr
    sizes, eps = ics.plummer_population(512, 1024, 131072, seed=7)
    pos_l, vel_l = [], []
    for k, n in enumerate(sizes):
        p, v = ics.plummer(int(n), 1.0, 0.3333, 1.0, seed=100000 + k)
        pos_l.append(p)
        vel_l.append(v)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    N = int(off[-1])
    pos = np.concatenate(pos_l, axis=1)
    vel = np.concatenate(vel_l, axis=1)
    mass = np.concatenate([np.full(n, 1.0 / n) for n in sizes])
    h = np.concatenate([np.full(n, 0.5 * e * e) for n, e in zip(sizes, eps)])
    sim = Sim(ctx, pos, vel, mass, h, 0.002)
    off_dev = torch.tensor(off, device="cuda:0", dtype=torch.int32)
    off_list = off.tolist()
    pairs = float(np.sum(sizes.astype(np.float64) ** 2))
    tune = None if retune else ctx.tune_batched_cached()
    tune_source = "cache"
    if tune is None:
        tune, _ = ctx.tune_batched(off_list)
        tune_source = "tuned"
    log(f"batched launch shape {tune.as_dict()} ({tune_source})")

    def grav_batched(out):
        # same composition and buffers every step: captured once, then replayed
        ctx.accel_batched(off_dev, off_list, sim.pos32, sim.h32, sim.m32, acc=out, tune=tune,
                          graph=True)

    # each system in its own buffers, as a one-system-per-process run holds it
    per_sys = []
    for s in range(len(sizes)):
        a, b = off_list[s], off_list[s + 1]
        per_sys.append((sim.pos32[:, a:b].contiguous(), sim.h32[a:b].contiguous(),
                        sim.m32[a:b].contiguous(), torch.zeros((3, b - a), device="cuda:0")))

    def grav_plain(out):
        del out
        for p, hh, mm, o in per_sys:
            ctx.accel(p, hh, p, mm, hh, acc=o)

    # the reference's one-sided method with the library's static launch shape
    one_sided = [ctx.tune_default(int(n)) for n in sizes]

    def grav_one_sided(out):
        del out
        for (p, hh, mm, o), t in zip(per_sys, one_sided):
            ctx.accel(p, hh, p, mm, hh, acc=o, tune=t)

    # the reference's launch shape: one body per lane, no j split
    ref_shape = khrr.Tune(1, 4, 1, 256, 0)

    def grav_ref_shape(out):
        del out
        for p, hh, mm, o in per_sys:
            ctx.accel(p, hh, p, mm, hh, acc=o, tune=ref_shape)

    def run(grav, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            grav(sim.acct)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k

    grav_batched(sim.acct)
    t_b = run(grav_batched, steps)
    log(f"batched: {t_b * 1e3:.1f} ms per force evaluation")
    grav_plain(sim.acct)
    t_p = run(grav_plain, max(1, steps // 2))
    log(f"plain per-system: {t_p * 1e3:.1f} ms per force evaluation")
    grav_one_sided(sim.acct)
    t_o = run(grav_one_sided, 1)
    log(f"plain per-system, one-sided: {t_o * 1e3:.1f} ms per force evaluation")
    grav_ref_shape(sim.acct)
    t_r = run(grav_ref_shape, 1)
    log(f"plain per-system, reference launch shape: {t_r * 1e3:.1f} ms per force evaluation")
    return {"systems": 512, "bodies": N, "pairs_per_step": pairs,
            "batched_ms": t_b * 1e3, "plain_ms": t_p * 1e3,
            "batched_Gpairs_per_s": pairs / t_b / 1e9, "plain_Gpairs_per_s": pairs / t_p / 1e9,
            "speedup_vs_plain": t_p / t_b, "tune": tune.as_dict(), "tune_source": tune_source,
            "graph": True,
            "plain": "khrr_accel per system, default launch shape, one stream",
            "plain_one_sided_ms": t_o * 1e3,
            "speedup_vs_plain_one_sided": t_o / t_b,
            "plain_one_sided": "khrr_accel per system with the static one-sided shape (every "
                               "directed pair evaluated, the reference's method)",
            "plain_reference_shape_ms": t_r * 1e3,
            "speedup_vs_plain_reference_shape": t_r / t_b,
            "plain_reference_shape": "khrr_accel per system with one body per lane and no "
                                     "j split (the reference's one-thread-per-body launch)"}


def bench_batched_run(ctx, steps, budget_gb=32.0):
    """The 512 Plummer spheres of the batched leg advanced `steps`
    predictor-corrector steps together: all submitted at once to one server,
    so every round is one segmented predictor launch, one batched gravity
    evaluation (a replayed graph) and one segmented corrector launch over
    all systems.  Reports the rate and each system's FP64 conservation."""
    import numpy as np
    import ics
    import khrr
    sizes, eps = ics.plummer_population(512, 1024, 131072, seed=7)
    grid = khrr.cube_grid(16, -2.0, 2.0)
    srv = khrr.Server(ctx, int(budget_gb * 1e9), len(sizes))
    ids = []
    for k, n in enumerate(sizes):
        p, v = ics.plummer(int(n), 1.0, 0.3333, 1.0, seed=100000 + k)
        ids.append(srv.submit(p, v, np.full(n, 1.0 / n), np.full(n, float(eps[k])), 0.002, steps,
                              grid))
    t0 = time.perf_counter()
    done = []
    while len(done) < len(ids):
        done += srv.step()
        st = srv.stats()
        if st["rounds"] % 100 == 0 and st["rounds"] > 0 and not done:
            log(f"batched run: round {st['rounds']}/{steps}")
    wall = time.perf_counter() - t0
    st = srv.stats()
    dE, dP, dL = [], [], []
    for rid in ids:
        d0, d1, _ = srv.result(rid, 16 ** 3)
        dE.append(abs(energy(d1) / energy(d0) - 1.0))
        dP.append(math.sqrt(sum((d1[k] - d0[k]) ** 2 for k in ("Px", "Py", "Pz"))))
        dL.append(math.sqrt(sum((d1[k] - d0[k]) ** 2 for k in ("Lx", "Ly", "Lz"))))
    srv.close()
    pairs = float(np.sum(sizes.astype(np.float64) ** 2)) * (steps + 3)
    dE, dP, dL = np.array(dE), np.array(dP), np.array(dL)
    return {"systems": len(sizes), "bodies": int(np.sum(sizes)), "steps": steps, "dt": 0.002,
            "rounds": st["rounds"], "wall_s": wall, "ms_per_step": wall / steps * 1e3,
            "Gpairs_per_s": pairs / wall / 1e9,
            "dE_rel_median": float(np.median(dE)), "dE_rel_max": float(np.max(dE)),
            "dP_abs_max": float(np.max(dP)), "dL_abs_max": float(np.max(dL)),
            "note": "per-system FP64 diagnostics at t = 0 and after the last step; "
                    "mass 1 per system, so dP and dL are absolute"}


def serving_specs(n_req, seed=7, tmax_lo=0.2, tmax_hi=2.0, dt=0.002):
    """The mixed request stream's shapes: 90 % Plummer spheres (N
    log-uniform on [1024, 131072], multiple of 1024), 10 % disk + halo
    galaxies (65536 + 131072); tmax uniform on [tmax_lo, tmax_hi].  Request
    r's bodies come from seed 200000 + r (request_bodies)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    specs = []
    for r in range(n_req):
        if rng.uniform() < 0.1:
            kind, n = "galaxy", 65536 + 131072
        else:
            kind = "plummer"
            n = int(np.clip(round(np.exp(rng.uniform(np.log(1024), np.log(131072))) / 1024) * 1024,
                            1024, 131072))
        steps = int(round(rng.uniform(tmax_lo, tmax_hi) / dt))
        specs.append(dict(r=r, kind=kind, n=n, steps=steps))
    return specs


def request_bodies(spec):
    """Bodies of one request: (pos, vel, mass, eps) host arrays.  Plummer
    spheres have mass 1, scale 0.3333, truncation 1.0 and eps 0.004 (N /
    16384)^-1/3; galaxies are the headline galaxy's model (eps 0.004)."""
    import numpy as np
    import ics
    seed = 200000 + spec["r"]
    if spec["kind"] == "galaxy":
        ps, vs, pd, vd = ics.galaxy(65536, 131072, seed=seed)
        pos, vel, mass, _ = ics.system_arrays(ps, vs, pd, vd, 1.0, 6.699356, 0.004, 0.004)
        return pos, vel, mass, np.full(pos.shape[1], 0.004)
    n = spec["n"]
    pos, vel = ics.plummer(n, 1.0, 0.3333, 1.0, seed=seed)
    return pos, vel, np.full(n, 1.0 / n), np.full(n, 0.004 * (n / 16384.0) ** (-1.0 / 3.0))


def serving_requests(n_req, seed=7, tmax_lo=0.2, tmax_hi=2.0, dt=0.002):
    """The first n_req requests of the stream with their bodies."""
    reqs = []
    for spec in serving_specs(n_req, seed, tmax_lo, tmax_hi, dt):
        pos, vel, mass, eps = request_bodies(spec)
        reqs.append(dict(pos=pos, vel=vel, mass=mass, eps=eps, steps=spec["steps"],
                         n=pos.shape[1]))
    return reqs


def poisson_arrivals(n_req, lam, seed=8):
    import numpy as np
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n_req))
    return arrivals - arrivals[0]


def serve_realtime(ctx, reqs, arrivals, budget_gb, max_active, round_pairs, grid, cells,
                   label=""):
    """Real-time arrivals into one server (request k submitted once
    arrivals[k] seconds have passed); returns the latencies and figures.
    busy: wall time while the server holds work (from the step that finds
    it non-empty to the step that drains it; a draining step finishes a
    request and so waits for the device)."""
    import numpy as np
    import khrr
    n_req = len(reqs)
    pairs = sum(r["n"] ** 2 * (r["steps"] + 3) for r in reqs)  # init acc+phi, steps, final phi
    srv = khrr.Server(ctx, int(budget_gb * 1e9), max_active, round_pairs=round_pairs)
    ctx.reset_timing()
    ctx.set_timing(True)
    ids, lat = {}, {}
    busy = 0.0
    busy_since = None
    t0 = time.perf_counter()
    last_log = t0
    nxt = 0
    while len(lat) < n_req:
        now = time.perf_counter() - t0
        while nxt < n_req and arrivals[nxt] <= now:
            r = reqs[nxt]
            ids[srv.submit(r["pos"], r["vel"], r["mass"], r["eps"], 0.002, r["steps"], grid)] = nxt
            nxt += 1
        st = srv.stats()
        if st["active"] == 0 and st["pending"] == 0:
            time.sleep(max(0.0, min(0.01, arrivals[nxt] - now)) if nxt < n_req else 0.0)
            continue
        if busy_since is None:
            busy_since = time.perf_counter()
        finished = srv.step()
        for fid in finished:
            k = ids[fid]
            lat[k] = time.perf_counter() - t0 - arrivals[k]
            srv.result(fid, cells)
        if n_req >= 50 and time.perf_counter() - last_log > 30.0:
            last_log = time.perf_counter()
            log(f"{label}served {len(lat)}/{n_req}, submitted {nxt}, active {st['active']}, "
                f"pending {st['pending']}, {last_log - t0:.0f} s")
        st = srv.stats()
        if st["active"] == 0 and st["pending"] == 0:
            busy += time.perf_counter() - busy_since
            busy_since = None
    makespan = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    _, grav_ms = ctx.kernel_time(khrr.K_GRAVITY)
    _, int_ms = ctx.kernel_time(khrr.K_INTEGRATE)
    ctx.set_timing(False)
    lat_s = np.array([lat[k] for k in range(n_req)])
    return lat_s, {"round_pairs": round_pairs, "makespan_s": makespan, "busy_s": busy,
                   "rounds": st["rounds"], "gravity_kernel_s": grav_ms * 1e-3,
                   "integrator_kernel_s": int_ms * 1e-3,
                   "pair_interactions": st["pair_interactions"],
                   "Gpairs_per_s": pairs / busy / 1e9,
                   "p50_latency_s": float(np.percentile(lat_s, 50)),
                   "p99_latency_s": float(np.percentile(lat_s, 99))}


def plain_service_times(ctx, reqs, grid, cells, label=""):
    """Each request alone with plain per-system launches (one khrr_accel per
    force evaluation, the step sequence of the drop-in), t = 0 and final
    diagnostics and the mass grid: its service time in seconds."""
    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    service = []
    last_log = time.perf_counter()
    for ri, r in enumerate(reqs):
        if len(reqs) >= 50 and time.perf_counter() - last_log > 30.0:
            last_log = time.perf_counter()
            log(f"{label}plain {ri}/{len(reqs)}")
        torch.cuda.synchronize()
        ts = time.perf_counter()
        sim = Sim(ctx, r["pos"], r["vel"], r["mass"], 0.5 * r["eps"] ** 2, 0.002)
        sim.init_forces()
        sim.diagnostics()
        for _ in range(r["steps"]):
            sim.step()
        sim.diagnostics()
        cnt = torch.zeros(cells, device=dev, dtype=torch.int32)
        cm = torch.zeros(cells, device=dev, dtype=torch.float64)
        ctx.p2mesh_bin(grid, sim.pos, sim.mass64, count=cnt, cell_mass=cm)
        cm.cpu()
        torch.cuda.synchronize()
        service.append(time.perf_counter() - ts)
        del sim
    return np.array(service)


def fifo_latency(arrivals, service):
    """Latencies of a FIFO queue serving one request at a time."""
    import numpy as np
    done = np.zeros(len(service))
    for k in range(len(service)):
        done[k] = max(arrivals[k], done[k - 1] if k else 0.0) + service[k]
    return done - np.asarray(arrivals[:len(service)])


def bench_serving(ctx, n_req, lam, budget_gb=8.0, max_active=64, round_pairs=3.4e10,
                  lam_saturated=50.0, tmax=(0.2, 2.0)):
    """Continuous batching on a Poisson arrival stream vs the same requests
    run one at a time (FIFO) with plain per-system launches; then the same
    requests arriving faster than either can serve them (lam_saturated)."""
    import numpy as np
    import khrr
    reqs = serving_requests(n_req, tmax_lo=tmax[0], tmax_hi=tmax[1])
    arrivals = poisson_arrivals(n_req, lam)
    grid = khrr.cube_grid(128, -2.0, 2.0)
    cells = 128 ** 3
    pairs = sum(r["n"] ** 2 * (r["steps"] + 3) for r in reqs)  # init acc+phi, steps, final phi
    policies = {}
    for name, rp in (("lockstep", 0.0), ("least_work_first", round_pairs)):
        lat_s, fig = serve_realtime(ctx, reqs, arrivals, budget_gb, max_active, rp, grid, cells)
        policies[name] = fig
        log(f"served ({name}) {n_req} requests in {fig['makespan_s']:.2f} s, "
            f"busy {fig['busy_s']:.2f} s, p50 {fig['p50_latency_s']:.3f} s, "
            f"p99 {fig['p99_latency_s']:.3f} s")
    service = plain_service_times(ctx, reqs, grid, cells)
    plain_lat = fifo_latency(arrivals, service)
    log(f"plain FIFO: busy {service.sum():.2f} s, p99 {np.percentile(plain_lat, 99):.3f} s")
    head = policies["least_work_first"]
    out = {"requests": n_req, "lambda_per_s": lam, "tmax_range": list(tmax), "dt": 0.002,
           "grid": "128^3 over [-2,2]^3", "budget_GB": budget_gb, "max_active": max_active,
           "mix": "90% Plummer N in [1024,131072] log-uniform, 10% galaxies 65536+131072",
           "pair_interactions": pairs,
           "policy": "least_work_first",
           "served_busy_s": head["busy_s"], "served_Gpairs_per_s": head["Gpairs_per_s"],
           "p50_latency_s": head["p50_latency_s"], "p99_latency_s": head["p99_latency_s"],
           "policies": policies,
           "plain_busy_s": float(service.sum()),
           "plain_Gpairs_per_s": pairs / service.sum() / 1e9,
           "plain_p50_latency_s": float(np.percentile(plain_lat, 50)),
           "plain_p99_latency_s": float(np.percentile(plain_lat, 99)),
           "throughput_vs_plain": float(service.sum() / head["busy_s"])}
    if lam_saturated <= 0:
        return out
    # saturated: every request queued within a fraction of a second
    arr_sat = poisson_arrivals(n_req, lam_saturated)
    lat_sat, fig_sat = serve_realtime(ctx, reqs, arr_sat, budget_gb, max_active, round_pairs,
                                      grid, cells)
    plain_sat = fifo_latency(arr_sat, service)
    saturated = {"lambda_per_s": lam_saturated, "policy": "least_work_first",
                 "served_busy_s": fig_sat["busy_s"], "served_Gpairs_per_s": fig_sat["Gpairs_per_s"],
                 "p50_latency_s": fig_sat["p50_latency_s"],
                 "p99_latency_s": fig_sat["p99_latency_s"], "rounds": fig_sat["rounds"],
                 "plain_p50_latency_s": float(np.percentile(plain_sat, 50)),
                 "plain_p99_latency_s": float(np.percentile(plain_sat, 99)),
                 "throughput_vs_plain": float(service.sum() / fig_sat["busy_s"])}
    log(f"served (saturated, lambda {lam_saturated}/s): busy {fig_sat['busy_s']:.2f} s, "
        f"p50 {fig_sat['p50_latency_s']:.3f} s, p99 {fig_sat['p99_latency_s']:.3f} s vs FIFO "
        f"p50 {saturated['plain_p50_latency_s']:.3f} s, p99 {saturated['plain_p99_latency_s']:.3f} s")
    out["saturated"] = saturated
    return out


def bench_serving_stream(ctx, n_req, lam, first, last, out_path, fifo=0, fifo_only=False,
                         budget_gb=8.0, max_active=64, round_pairs=3.4e10, tmax=(0.2, 2.0)):
    """Requests [first, last) of the n_req-request serving stream through one
    server in real time, each submitted at its Poisson arrival offset from
    request `first` (lam arrivals per second over the whole stream), with
    the least-remaining-work round schedule.  The whole stream is longer
    than one run may take, so it runs as consecutive segments whose records
    --stream-merge combines; requests still running at a segment's end
    finish in it, so the next segment starts with an empty server.  fifo >
    0: the segment's first `fifo` requests are also run one at a time with
    plain launches, for the FIFO latencies of that prefix; fifo_only: only
    that (no served run).  Writes per-request records (arrival, latency,
    size, steps) to out_path."""
    import numpy as np
    import khrr
    specs = serving_specs(n_req, tmax_lo=tmax[0], tmax_hi=tmax[1])[first:last]
    arrivals = poisson_arrivals(n_req, lam)
    log(f"stream: generating requests {first}..{last - 1}")
    reqs = []
    for spec in specs:
        pos, vel, mass, eps = request_bodies(spec)
        reqs.append(dict(pos=pos, vel=vel, mass=mass, eps=eps, steps=spec["steps"],
                         n=pos.shape[1]))
    arr = arrivals[first:last] - arrivals[first]
    grid = khrr.cube_grid(128, -2.0, 2.0)
    cells = 128 ** 3
    rec = {"stream_requests": n_req, "lambda_per_s": lam, "tmax_range": list(tmax), "dt": 0.002,
           "grid": "128^3 over [-2,2]^3", "budget_GB": budget_gb, "max_active": max_active,
           "policy": "least_work_first", "round_pairs": round_pairs,
           "first": first, "last": last, "fifo_only": bool(fifo_only),
           "r": [s["r"] for s in specs], "kind": [s["kind"] for s in specs],
           "n": [s["n"] for s in specs], "steps": [s["steps"] for s in specs],
           "arrival_s": [float(a) for a in arrivals[first:last]]}
    if not fifo_only:
        log(f"stream: serving {len(reqs)} requests over {arr[-1]:.0f} s of arrivals")
        lat, fig = serve_realtime(ctx, reqs, arr, budget_gb, max_active, round_pairs, grid,
                                  cells, label="stream: ")
        log(f"stream: busy {fig['busy_s']:.1f} s, {fig['Gpairs_per_s']:.0f} Gpairs/s, "
            f"p50 {fig['p50_latency_s']:.2f} s, p99 {fig['p99_latency_s']:.2f} s")
        rec["figures"] = fig
        rec["latency_s"] = [float(x) for x in lat]
    else:
        fifo = len(reqs)
    if fifo > 0:
        k = min(fifo, len(reqs))
        service = plain_service_times(ctx, reqs[:k], grid, cells, label="stream: ")
        rec["fifo_service_s"] = [float(x) for x in service]
        rec["fifo_latency_s"] = [float(x) for x in fifo_latency(arr[:k], service)]
        log(f"stream: plain FIFO over the first {k}: busy {service.sum():.1f} s")
    with open(out_path, "w") as f:
        json.dump(rec, f)
    return {k: v for k, v in rec.items() if not isinstance(v, list) or k == "tmax_range"}


def stream_merge(paths):
    """Combines the records of consecutive serving-stream segments: latency
    percentiles over every request, aggregate pair-interactions per second
    over the segments' busy time, and the FIFO comparison on the prefix."""
    import numpy as np
    allrecs = [json.load(open(p)) for p in paths]
    recs = sorted((r for r in allrecs if not r.get("fifo_only")), key=lambda r: r["first"])
    for a, b in zip(recs, recs[1:]):
        if a["last"] != b["first"]:
            raise SystemExit(f"segments not contiguous: {a['last']} != {b['first']}")
    lat = np.concatenate([r["latency_s"] for r in recs])
    n = np.concatenate([r["n"] for r in recs]).astype(np.float64)
    steps = np.concatenate([r["steps"] for r in recs]).astype(np.float64)
    pairs = float(np.sum(n * n * (steps + 3)))
    busy = sum(r["figures"]["busy_s"] for r in recs)
    span = sum(r["figures"]["makespan_s"] for r in recs)
    gal = np.concatenate([np.array(r["kind"]) == "galaxy" for r in recs])
    out = {"requests": int(len(lat)), "first": recs[0]["first"], "last": recs[-1]["last"],
           "segments": [[r["first"], r["last"]] for r in recs],
           "lambda_per_s": recs[0]["lambda_per_s"], "tmax_range": recs[0]["tmax_range"],
           "budget_GB": recs[0]["budget_GB"], "max_active": recs[0]["max_active"],
           "grid": recs[0]["grid"], "policy": recs[0]["policy"],
           "galaxy_requests": int(gal.sum()), "pair_interactions": pairs,
           "busy_s": busy, "makespan_s": span, "Gpairs_per_s_busy": pairs / busy / 1e9,
           "p50_latency_s": float(np.percentile(lat, 50)),
           "p90_latency_s": float(np.percentile(lat, 90)),
           "p99_latency_s": float(np.percentile(lat, 99)),
           "max_latency_s": float(np.max(lat)),
           "p50_latency_galaxy_s": float(np.percentile(lat[gal], 50)) if gal.any() else None,
           "p99_latency_galaxy_s": float(np.percentile(lat[gal], 99)) if gal.any() else None}
    fifo = [r for r in allrecs if "fifo_latency_s" in r]
    if fifo:
        r = fifo[0]
        k = len(r["fifo_latency_s"])
        # served latencies of the same requests, from the segment records
        by_r = {rr: lt for seg in recs for rr, lt in zip(seg["r"], seg["latency_s"])}
        if not all(rr in by_r for rr in r["r"][:k]):
            raise SystemExit("the FIFO requests are not all in the served segments")
        served = np.array([by_r[rr] for rr in r["r"][:k]])
        fl = np.array(r["fifo_latency_s"])
        sv = np.array(r["fifo_service_s"])
        nn = np.array(r["n"][:k], np.float64)
        st = np.array(r["steps"][:k], np.float64)
        out["fifo_prefix"] = {
            "requests": [r["first"], r["first"] + k],
            "plain_busy_s": float(sv.sum()),
            "plain_Gpairs_per_s": float(np.sum(nn * nn * (st + 3)) / sv.sum() / 1e9),
            "plain_fifo_p50_latency_s": float(np.percentile(fl, 50)),
            "plain_fifo_p99_latency_s": float(np.percentile(fl, 99)),
            "served_p50_latency_s": float(np.percentile(served, 50)),
            "served_p99_latency_s": float(np.percentile(served, 99))}
    return out


def host_cpu_info():
    """CPU model, physical cores and logical CPUs of this host from
    /proc/cpuinfo (physical cores: distinct (physical id, core id) pairs)."""
    model, cores, logical = "", set(), 0
    ident = {}
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for ln in f:
                key, _, val = ln.partition(":")
                key, val = key.strip(), val.strip()
                if key == "processor":
                    logical += 1
                elif key == "model name" and not model:
                    model = val
                elif key in ("vendor_id", "cpu family", "model", "stepping", "cpu MHz") and \
                        key not in ident:
                    ident[key] = val
                elif key == "flags" and "isa" not in ident:
                    have = set(val.split())
                    ident["isa"] = " ".join(x for x in ("avx2", "fma", "avx512f", "avx512dq",
                                                         "avx512_bf16") if x in have)
                elif key == "physical id":
                    phys = val
                elif key == "core id":
                    core = val
                elif not key and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
            if phys is not None:
                cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "cpu_id": ident, "physical_cores": len(cores) or None,
            "logical_cpus": logical or None, "affinity_cpus": len(os.sched_getaffinity(0))}


def median_time(fn, reps=3):
    """Median wall time of reps calls of fn (and every run's time)."""
    import numpy as np
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_leg(checks, rows_n, result):
    """The CPU leg: the FP64 oracle (OpenMP, the host cores given to this
    process) timed on bounded samples of the headline workload -- the
    reported CPU baseline, median of 3 runs each: gravity rows, binning on
    the headline grid, stencil iterations and the 24576-body galaxy's step
    loop -- and, as the checker of the GPU legs, the FP64 references behind
    their error figures: headline forces at the final positions, the
    stencil field after 200 iterations, the precision form's forces."""
    import numpy as np
    import ics
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    hd = checks["headline"]
    n = hd["pos"].shape[1]
    rows = np.sort(np.random.default_rng(4096).choice(n, rows_n, replace=False))
    out = {}

    def gravity_rows():
        out["ref"], _ = orc.accel_rows(rows, hd["pos"], hd["h"], hd["m"])

    dt, dts = median_time(gravity_rows)
    ref = out["ref"]
    e = np.linalg.norm(hd["acc"][:, rows] - ref, axis=0) / np.linalg.norm(ref, axis=0)
    result["accuracy"] = {"rows": int(rows_n), "median": float(np.median(e)),
                          "p99": float(np.percentile(e, 99)), "max": float(np.max(e)),
                          "reference": "FP64 oracle on the same FP32 inputs"}
    threads = orc.num_threads()
    cpu = {"value": rows_n * n / dt / 1e9, "unit": "G pair-interactions/s",
           "cores": threads, "kind": "port",
           "sample": f"{rows_n} random i-rows x {n} j of the headline galaxy (a force "
                     f"evaluation subsample, FP64, OpenMP {threads} threads), median of 3 "
                     f"runs: {dt:.2f} s",
           "gravity_runs_s": dts}
    cpu.update(host_cpu_info())
    log(f"cpu leg: gravity {cpu['value']:.2f} G pairs/s on {threads} threads")
    # binning of the headline galaxy's final state on 256^3 over [-2,2]^3
    bn = checks.get("binning")
    if bn is not None:
        nx = bn["nx"]

        def bin_oracle():
            out["bins"] = orc.p2mesh_bin([-2.0] * 3, [4.0 / nx] * 3, [nx] * 3, bn["pos"],
                                         bn["mass"])

        tb, tbs = median_time(bin_oracle)
        exact = all(np.array_equal(bn["out"][k], out["bins"][k]) for k in bn["out"])
        result["binning"]["bit_exact_vs_oracle"] = bool(exact)
        log(f"cpu leg: binning outputs bit-exact vs the oracle: {exact}")
        cpu["binning_ms"] = tb * 1e3
        cpu["binning_runs_ms"] = [t * 1e3 for t in tbs]
        cpu["binning_sample"] = (f"{bn['pos'].shape[1]} bodies on {nx}^3 cells, every output; "
                                 "serial (index-order ranks and sums), median of 3")
        log(f"cpu leg: binning {tb * 1e3:.1f} ms")
    # stencil iterations on the headline grid (OpenMP)
    st = checks.get("stencil")
    if st is not None:
        nx, k = st["nx"], 10
        G = nx ** 3
        w = orc.reference_wave(nx, 4.0 / nx)
        ts, tss = median_time(lambda: orc.wave_iterate(w, np.zeros(G), np.zeros(G),
                                                       st["cell_mass"], k))
        cpu["stencil_ms_per_iter"] = ts / k * 1e3
        cpu["stencil_runs_ms_per_iter"] = [t / k * 1e3 for t in tss]
        cpu["stencil_sample"] = (f"{nx}^3, {k} iterations per run from phi = 0, FP64, OpenMP "
                                 f"{threads} threads, median of 3")
        log(f"cpu leg: stencil {ts / k * 1e3:.1f} ms per iteration")
    # the 8192 + 16384-body galaxy's predictor -> ACCEL -> corrector loop
    ps, vs, pd, vd = ics.galaxy(8192, 16384, seed=12345)
    p0, v0, m0, h0 = ics.system_arrays(ps, vs, pd, vd, 1.0, 6.699356, 0.004, 0.004)
    h0 = h0.astype(np.float32).astype(np.float64)
    c0_steps = 25
    tc, tcs = median_time(lambda: orc.run(p0, v0, m0, h0, 0.002, c0_steps, with_diag=False))
    cpu["c0_500_steps_s"] = tc / c0_steps * 500
    cpu["c0_runs_s"] = tcs
    cpu["c0_sample"] = (f"24576-body galaxy, {c0_steps} predictor-ACCEL-corrector steps per "
                        f"run (FP64, OpenMP {threads} threads), median of 3, scaled to 500 "
                        "steps (every step costs the same)")
    log(f"cpu leg: 24576-body galaxy {tc / c0_steps * 1e3:.0f} ms per step")
    result["cpu_baseline"] = cpu
    st = checks.get("stencil")
    if st and "stencil" in result:
        log("cpu leg: FP64 stencil reference")
        nx = st["nx"]
        ref_phi, _ = orc.wave_iterate(orc.reference_wave(nx, 4.0 / nx), np.zeros(nx ** 3),
                                      np.zeros(nx ** 3), st["cell_mass"], st["iters"])
        result["stencil"]["err_vs_fp64"] = float(np.max(np.abs(st["phi"] - ref_phi)) /
                                                 np.max(np.abs(ref_phi)))
        result["stencil"]["err_definition"] = "max|phi32 - phi64| / max|phi64| after all iterations"
    pr = checks.get("precision")
    if pr and "precision" in result:
        log("cpu leg: FP64 forces of the precision-form sample")
        ref, _ = orc.accel_rows(pr["rows"], pr["pos"], pr["h"], pr["m"])
        for mode, acc in pr["acc"].items():
            e = np.linalg.norm(acc - ref, axis=0) / np.linalg.norm(ref, axis=0)
            result["precision"][mode].update(err_median=float(np.median(e)),
                                             err_p99=float(np.percentile(e, 99)),
                                             err_max=float(np.max(e)))
            log(f"precision {mode}: error median {np.median(e):.2e} "
                f"p99 {np.percentile(e, 99):.2e}")


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path is built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.stream_merge:
        print(json.dumps(stream_merge(args.stream_merge.split(","))), flush=True)
        return
    import numpy as np
    import torch
    import ics
    import khrr

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.stream_range:
        first, last = (int(v) for v in args.stream_range.split(","))
        if not 0 <= first < last <= args.stream_requests:
            raise SystemExit(f"--stream-range {args.stream_range} outside the stream")
        ctx = khrr.Context(0)
        if args.tune_cache and os.path.exists(args.tune_cache):
            try:
                ctx.tune_load(args.tune_cache)
            except khrr.KhrrError as e:
                log(f"tune cache not used: {e}")
        tmax = tuple(float(v) for v in args.serve_tmax.split(","))
        out = bench_serving_stream(ctx, args.stream_requests, args.serve_lambda, first, last,
                                   args.stream_out, fifo=args.stream_fifo,
                                   fifo_only=args.stream_fifo_only, tmax=tmax,
                                   round_pairs=args.serve_round_pairs)
        print(json.dumps(out), flush=True)
        return
    if args.calib_copy:
        # a plain copy of known size: 512 MiB read + 512 MiB written
        a = torch.ones(128 * 1024 * 1024, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("calibration copy: 536870912 bytes read + written")
    log("generating galaxy")
    ps, vs, pd, vd = ics.galaxy(args.n_stars, args.n_dm, seed=12345)
    pos, vel, mass, h = ics.system_arrays(ps, vs, pd, vd, 1.0, 6.699356, 0.004, 0.004)
    n = pos.shape[1]
    ctx = khrr.Context(0)
    # launch shape from the project's autotuner (cached in the context for
    # this size bucket; every later khrr_accel of this size uses it)
    tune_source = "fixed"
    if args.tune_cache and os.path.exists(args.tune_cache) and not args.retune:
        try:
            ctx.tune_load(args.tune_cache)
            log(f"tune cache {os.path.relpath(args.tune_cache, ROOT)} loaded")
        except khrr.KhrrError as e:
            log(f"tune cache not used: {e}")
    if args.shape:
        tune = khrr.Tune(*[int(v) for v in args.shape.split(",")])
        log(f"fixed launch shape {tune.as_dict()}")
    else:
        tune = None if args.retune else ctx.tune_cached(n)
        tune_source = "cache"
        if tune is None:
            tune, tune_ms = ctx.tune_gravity(n)
            tune_source = "tuned"
            log(f"autotuned launch shape {tune.as_dict()} ({tune_ms:.2f} ms per trial)")
        else:
            log(f"launch shape {tune.as_dict()} from the tune cache")
    sim = Sim(ctx, pos, vel, mass, h, 0.002)
    sim.tune = tune
    sim.init_forces()
    torch.cuda.synchronize()
    log(f"n={n}; warmup {args.warmup} steps")
    d0 = sim.diagnostics()
    for _ in range(args.warmup):
        sim.step()
    torch.cuda.synchronize()
    d_start = sim.diagnostics()
    torch.cuda.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    grav_launches, grav_ms = ctx.kernel_time(khrr.K_GRAVITY)
    int_launches, int_ms = ctx.kernel_time(khrr.K_INTEGRATE)
    # the conservation-law reductions, timed on their own
    ctx.reset_timing()
    for _ in range(10):
        sim.diagnostics()
    diag_calls, diag_ms = ctx.kernel_time(khrr.K_DIAG)
    ctx.set_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    pairs = float(n) * n
    value = pairs / (ms_per_step * 1e-3) / 1e9
    grav_avg = grav_ms / max(grav_launches, 1)
    achieved = FLOP_PER_PAIR * pairs / (grav_avg * 1e-3) / 1e12
    if tune.symmetric:
        # off-diagonal superblock pairs once, diagonal superblocks one-sided
        # (the shape's superblock size, else the library's rule of n)
        sb = tune.superblock or (4096 if n >= 524288 else (2048 if n >= 32768 else 1024))
        sizes = [min(sb, n - a) for a in range(0, n, sb)]
        diag_pairs = float(sum(s * s for s in sizes))
        executed = (pairs - diag_pairs) / 2 + diag_pairs
    else:
        executed = pairs
    log(f"{ms_per_step:.2f} ms/step, gravity {grav_avg:.2f} ms/launch, {value:.1f} Gpairs/s")
    one_sided = None
    if tune.symmetric:
        # the one-sided kernel on the same bodies (every directed pair
        # evaluated), for the frac of a kernel that executes all N^2 pairs
        os_tune = khrr.Tune(2, 4, 8, 256, 0)
        sim.tune = os_tune
        out = torch.zeros((3, n), device="cuda:0")
        sim.gravity(out)
        torch.cuda.synchronize()
        ctx.reset_timing()
        ctx.set_timing(True)
        for _ in range(2):
            sim.gravity(out)
        os_l, os_ms = ctx.kernel_time(khrr.K_GRAVITY)
        ctx.set_timing(False)
        sim.tune = tune
        del out
        os_avg = os_ms / max(os_l, 1)
        os_ach = FLOP_PER_PAIR * pairs / (os_avg * 1e-3) / 1e12
        one_sided = {"kernel": "grav_kernel", "launch_shape": os_tune.as_dict(),
                     "avg_launch_ms": os_avg, "achieved": os_ach,
                     "frac": os_ach / PEAK_FP32_TFLOPS,
                     "G_pair_interactions_per_s": pairs / (os_avg * 1e-3) / 1e9}
        log(f"one-sided kernel: {os_avg:.2f} ms/launch, frac {os_ach / PEAK_FP32_TFLOPS:.3f}")
    d_end = sim.diagnostics()
    drift = {"steps": args.steps,
             "dE_rel": energy(d_end) / energy(d_start) - 1.0,
             "dP_abs": math.sqrt(sum((d_end[k] - d_start[k]) ** 2 for k in ("Px", "Py", "Pz"))),
             "dLz_rel": d_end["Lz"] / d_start["Lz"] - 1.0,
             "E0": energy(d0)}
    # HBM bytes per launch from the committed counter passes, when they were
    # taken on the kernel and launch shape this run uses
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            tf = json.load(open(TRAFFIC_FILE))
            shape = tf.get("launch_shape", {})
            same = (tf.get("kernel", "").startswith("sym_kernel" if tune.symmetric else "grav_kernel")
                    and all(getattr(tune, k) == v for k, v in shape.items()))
            if same:
                traffic = tf.get("gravity_bytes_per_launch")
        except (OSError, ValueError, AttributeError):
            traffic = None
    result = {
        "metric": METRIC, "value": value, "unit": "G pair-interactions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "disk+halo galaxy, 262144 stars + 524288 halo, eps 0.004, "
                               "dt 0.002, FP32 direct sum, FP64 predictor-corrector state",
                   "n_bodies": n, "dt": 0.002, "seed": 12345,
                   "launch_shape": tune.as_dict(), "tune_source": tune_source},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic,
                     "kernel": "sym_kernel" if tune.symmetric else "grav_kernel",
                     "avg_launch_ms": grav_avg, "flop_per_pair": FLOP_PER_PAIR,
                     "pairs_per_launch": pairs,
                     "executed_pair_evaluations_per_launch": executed,
                     "note": ("pair-interactions counted as N^2 directed interactions per "
                              "force evaluation (the reference's i = j included); the "
                              "pair-symmetric kernel evaluates each unordered pair once and "
                              "applies it to both bodies, so it executes "
                              f"{executed / pairs:.4f} N^2 pair evaluations"),
                     "one_sided": one_sided},
        "integrator_ms_per_launch": int_ms / max(int_launches, 1),
        # algorithmic bytes per body: predictor 120 (reads pos, vel, acc;
        # writes pos_t, vel_t and the FP32 position copy), corrector 144
        # (reads pos_t, vel_t, vel, acc; writes pos, vel and the force
        # copy), 132 on average over the two launches of a step
        "integrator": hbm_figure(int_ms / max(int_launches, 1), 132 * n),
        # FP64 reductions of khrr_diag (its two kernels): read pos, vel,
        # mass, phi, 60 B per body
        "diagnostics": hbm_figure(diag_ms / max(diag_calls, 1), 60 * n),
        "drift": drift,
    }
    # what the CPU leg checks, collected by the GPU legs (host copies)
    checks = {}
    if not args.no_cpu:
        # fresh force at the final positions for the accuracy sample
        chk = torch.zeros((3, n), device="cuda:0")
        sim.gravity(chk)
        torch.cuda.synchronize()
        checks["headline"] = dict(pos=sim.pos32.cpu().numpy().astype(np.float64),
                                  m=sim.m32.cpu().numpy().astype(np.float64),
                                  h=sim.h32.cpu().numpy().astype(np.float64),
                                  acc=chk.cpu().numpy().astype(np.float64))
        del chk
    legs = set() if args.quick else set(args.legs.split(","))
    if legs:
        ctx.set_timing(True)
    if legs & {"binning", "stencil", "mesh"}:
        log("binning")
        bins, binning = bench_binning(ctx, sim)
        result["binning"] = binning
        if not args.no_cpu:  # the CPU leg times the oracle on it and checks the outputs
            checks["binning"] = dict(nx=256, pos=sim.pos.cpu().numpy(),
                                     mass=sim.mass64.cpu().numpy(),
                                     out={k: v.cpu().numpy() for k, v in bins.items()})
        if legs & {"stencil", "mesh"}:
            log("stencil")
            checks["stencil"] = {}
            field, result["stencil"] = bench_stencil(ctx, bins["cell_mass"], checks["stencil"])
            if "mesh" in legs:
                log("mesh-side entry points")
                result["mesh"] = bench_mesh(ctx, sim, bins, field)
            del field
        del bins
        torch.cuda.empty_cache()
    if "batched" in legs:
        log("batched")
        result["batched"] = bench_batched(ctx, args.batched_steps, retune=args.retune)
        torch.cuda.empty_cache()
    if "batched-run" in legs:
        log(f"batched run: 512 systems, {args.batched_run_steps} steps")
        result["batched_run"] = bench_batched_run(ctx, args.batched_run_steps)
        torch.cuda.empty_cache()
    if "serving" in legs and args.serve_requests > 0:
        log("serving")
        tmax = tuple(float(v) for v in args.serve_tmax.split(","))
        result["serving"] = bench_serving(ctx, args.serve_requests, args.serve_lambda,
                                          lam_saturated=args.serve_saturated, tmax=tmax,
                                          round_pairs=args.serve_round_pairs)
        torch.cuda.empty_cache()
    if "precision" in legs and args.precision_steps > 0:
        log("precision form")
        checks["precision"] = {}
        result["precision"] = bench_precision(ctx, args.precision_steps, checks["precision"])
    if not args.no_cpu:
        log(f"cpu leg: oracle on {args.cpu_rows} rows of the headline workload")
        cpu_leg(checks, args.cpu_rows, result)
    if args.tune_save:
        ctx.tune_save(args.tune_save)
        log(f"tuned shapes written to {args.tune_save}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
