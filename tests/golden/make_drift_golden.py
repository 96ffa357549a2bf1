"""Writes tests/golden/galaxy24k_drift_golden.npz: the FP64 oracle's run of the
small plain-method workload, for the drift comparison of the GPU step loop
(run once, in the build container; the output is committed, the GPU host
never runs this).

Workload: one galaxy without external potential, 8192 disk stars (mass 1.0)
and 16384 Plummer halo bodies (mass 6.699356), eps 0.004, dt 0.002, seed
12345, 500 predictor-corrector steps (the reference's step loop
nbody.cu:1518-1582 with the force accumulator cleared every step).
The fixture holds the 17 diagnostics (Ek, Ep, P, L, M, M*r, maxima; order
of result() nbody.cu:797-833) at t = 0 and after every 10th step, plus
checksums of the initial conditions so the test can confirm it rebuilds the
same bodies.

Usage: python tests/golden/make_drift_golden.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "khrr-nbody_amd"))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

N_S, N_DM, SEED, DT, STEPS, EVERY = 8192, 16384, 12345, 0.002, 500, 10


def workload():
    import ics
    ps, vs, pd, vd = ics.galaxy(N_S, N_DM, seed=SEED)
    return ics.system_arrays(ps, vs, pd, vd, 1.0, 6.699356, 0.004, 0.004)


def main():
    import oracle as orc
    pos, vel, mass, h = workload()
    # the GPU path holds mass and softening in FP32; the oracle gets the same
    # values widened, positions and velocities stay FP64 in both
    m32 = mass.astype(np.float32).astype(np.float64)
    h32 = h.astype(np.float32).astype(np.float64)
    t0 = time.perf_counter()
    _, _, d = orc.run(pos, vel, m32, h32, DT, STEPS, diag_every=EVERY)
    print(f"oracle: {STEPS} steps of n={pos.shape[1]} in {time.perf_counter() - t0:.1f} s")
    np.savez_compressed(os.path.join(HERE, "galaxy24k_drift_golden.npz"), diag=d, dt=DT, steps=STEPS,
                        every=EVERY, n_s=N_S, n_dm=N_DM, seed=SEED,
                        ic_checksum=np.array([pos.sum(), vel.sum(), (pos * pos).sum(),
                                              (vel * vel).sum(), m32.sum(), h32.sum()]))


if __name__ == "__main__":
    main()
