"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np


def norm_err(gpu, ref, mask=None) -> float:
    """max_i |gpu_i - ref_i| / max_i |ref_i| over the (masked) particles; vectors use the Euclidean norm."""
    g = np.asarray(gpu, np.float64)
    r = np.asarray(ref, np.float64)
    if mask is not None:
        g, r = g[mask], r[mask]
    if g.size == 0:
        return 0.0
    if g.ndim == 2:
        d = np.linalg.norm(g - r, axis=1)
        s = np.linalg.norm(r, axis=1)
    else:
        d = np.abs(g - r)
        s = np.abs(r)
    scale = s.max()
    if scale == 0.0:
        return float(d.max())
    return float(d.max() / scale)


def report(name, errs: dict) -> str:
    return name + ": " + ", ".join(f"{k}={v:.2e}" for k, v in errs.items())
