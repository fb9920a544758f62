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


def ghost_gain(bundle, Nxx, **params):
    """Largest sum of |coefficients| over the rows of the composed ghost map: a ghost value is
    that linear combination of interior values, so a difference between two states' interior
    values reaches the ghost points multiplied by at most this factor (31^3 at a corner of
    three outermost layers of quadratic extrapolation: 10 f0 - 15 f1 + 6 f2 per direction)."""
    from nrpytutorial_amd.runtime import HostBC
    _, rowptr, _, coeff, _ = HostBC(bundle, Nxx, **params).composed()
    sums = np.add.reduceat(np.abs(coeff), rowptr[:-1]) if len(coeff) else np.zeros(1)
    return float(sums.max())


def interior_ghost_err(a, b, G, gain, report=False):
    """(interior error, ghost error / gain), each the per-gridfunction max|a-b| relative to
    max|b| over interior points, maximised over gridfunctions. Two evolutions that differ by
    rounding in the interior differ at ghost points by up to ``gain`` times that (plus the
    fill's own rounding), so both numbers are held to the same tolerance."""
    if not (np.isfinite(a).all() and np.isfinite(b).all()):
        return float("inf"), float("inf")
    inner = (slice(None), slice(G, -G), slice(G, -G), slice(G, -G))
    mask = np.ones(a.shape[1:], bool)
    mask[G:-G, G:-G, G:-G] = False
    worst_i = worst_g = 0.0
    for q in range(a.shape[0]):
        scale = float(np.abs(b[inner][q]).max()) or 1.0
        ei = float(np.abs(a[inner][q] - b[inner][q]).max()) / scale
        eg = float(np.abs(a[q][mask] - b[q][mask]).max()) / scale / gain
        if report:
            print("  gf %2d  interior rel %.3e  ghost rel/gain %.3e  (scale %.3e)" % (q, ei, eg, scale))
        worst_i, worst_g = max(worst_i, ei), max(worst_g, eg)
    return worst_i, worst_g


def per_gf_rel(a, b):
    """max|a-b| / max|b| per gridfunction (interior arrays)."""
    return np.array([np.abs(a[q] - b[q]).max() / max(np.abs(b[q]).max(), 1e-300) for q in range(a.shape[0])])


def closer_to_exact(got, ref, exact, factor=2.0, floor=1e-13):
    """The extended-precision criterion: per gridfunction, |got - exact| within ``factor``
    times the FP64 oracle's own |ref - exact|, or below ``floor`` where both are at rounding
    level (errors relative to each gridfunction's max |exact|). Returns (ok, e_got, e_ref)."""
    e_got, e_ref = per_gf_rel(got, exact), per_gf_rel(ref, exact)
    return bool(np.all(e_got <= np.maximum(factor * e_ref, floor))), e_got, e_ref
