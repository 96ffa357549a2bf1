"""Small shared helpers for the tests (inputs and error statistics)."""
import numpy as np

import ics


def galaxy_system(n_s, n_dm, seed=12345, eps_s=0.004, eps_d=0.004):
    """Disk + halo galaxy in reference particle order: (pos, vel, mass, h)
    as float64 arrays (h = eps^2/2)."""
    ps, vs, pd, vd = ics.galaxy(n_s, n_dm, seed=seed)
    return ics.system_arrays(ps, vs, pd, vd, 1.0, 6.699356, eps_s, eps_d)


def rel_err(a, ref):
    """Per-particle |a - ref| / |ref| over the 3 components (columns)."""
    num = np.linalg.norm(np.asarray(a, np.float64) - ref, axis=0)
    den = np.linalg.norm(ref, axis=0)
    return num / den


def err_stats(e):
    return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
            "max": float(np.max(e))}


def to_dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda:0", dtype=dtype).contiguous()
