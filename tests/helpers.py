"""Shared helpers for the parity tests."""
import math

import numpy as np


def splitmix64(seed):
    """Deterministic 64-bit generator (Steele, Lea & Flood 2014), used to draw
    documented synthetic inputs independent of numpy's generator versions."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def uniform(seed, n):
    g = splitmix64(seed)
    return [(next(g) >> 11) * (1.0 / 9007199254740992.0) for _ in range(n)]


def unit_direction(seed):
    """Unit vector from 3 standard normals (Box-Muller on splitmix64 uniforms)."""
    u = uniform(seed, 6)
    v = []
    for k in range(3):
        r = math.sqrt(-2.0 * math.log(1.0 - u[2 * k]))
        v.append(r * math.cos(2.0 * math.pi * u[2 * k + 1]))
    n = math.sqrt(sum(c * c for c in v))
    return [c / n for c in v]


def rel_maxnorm_err(a, b, floor_frac=1e-12, report=False):
    """Per-gridfunction max|a-b| / max|b|, maximised over gridfunctions.

    Gridfunctions that vanish analytically (e.g. off-axis components in
    axisymmetry) hold only roundoff noise; their scale is floored at
    ``floor_frac`` times the largest max-norm of all gridfunctions, so the
    measure is relative to the state's magnitude for them."""
    if np.isnan(a).any() or np.isnan(b).any():
        return float("inf")
    gmax = float(np.max(np.abs(b))) if b.size else 0.0
    worst = 0.0
    rows = []
    for g in range(a.shape[0]):
        scale = max(float(np.max(np.abs(b[g]))), floor_frac * gmax)
        d = float(np.max(np.abs(a[g] - b[g])))
        e = d / scale if scale > 0 else d
        rows.append((g, d, scale, e))
        worst = max(worst, e)
    if report:
        for g, d, s, e in rows:
            print("  gf %2d  max|diff| %.3e  scale %.3e  rel %.3e" % (g, d, s, e))
    return worst
