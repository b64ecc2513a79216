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


def point_err(gpu, ref, mask=None, floor=1e-3) -> float:
    """max_i |gpu_i - ref_i| / max(|ref_i|, floor * max_j |ref_j|) over the
    (masked) particles: the error relative to each particle's own value, a
    value below `floor` of the largest being measured against that floor
    (vectors use the Euclidean norm)."""
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
    den = np.maximum(s, floor * s.max())
    den[den == 0.0] = 1.0
    return float((d / den).max())


def report(name, errs: dict) -> str:
    return name + ": " + ", ".join(f"{k}={v:.2e}" for k, v in errs.items())
