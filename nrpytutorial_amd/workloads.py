"""Seeded synthetic inputs for the benchmark and parity workloads.

Random numbers come from splitmix64 (Steele, Lea & Flood, 2014) so inputs are
reproducible across numpy/torch versions.

* ``plane_wave_direction(seed)``: unit vector from three Box-Muller normals
  (scalar-wave plane wave uu = 2 + sin(k.x), vv = -cos(k.x)).
* ``gaussian_bumps(seed, ngf)``: per gridfunction, 8 Gaussians
  A exp(-|x - c|^2 / w^2) with A = 1e-3, w in U[1,3], c in U[-10,10]^3, added
  to the flat BSSN state (h=a=lambda=vet=bet=trK=0, cf=alpha=1): the
  "flat metric plus a sum of 8 seeded Gaussians in every field" input of the
  192^3 BSSN RHS benchmark.
* ``scalarwave_bump_parameters`` / ``scalarwave_bump_basis_*``: the adjoint
  scalar-wave input uu = 2 + sum_k a_k exp(-|x - c_k|^2 / (2 * 1.5^2)), vv = 0,
  16 bumps with centres in [-5,5]^3.
"""
import math

import numpy as np

BSSN_FLAT_ONE = ("cf", "alpha")


def splitmix64(seed):
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def uniforms(seed, n):
    g = splitmix64(seed)
    return [(next(g) >> 11) * (1.0 / 9007199254740992.0) for _ in range(n)]


def plane_wave_direction(seed):
    u = uniforms(seed, 6)
    v = []
    for k in range(3):
        r = math.sqrt(-2.0 * math.log(1.0 - u[2 * k]))
        v.append(r * math.cos(2.0 * math.pi * u[2 * k + 1]))
    n = math.sqrt(sum(c * c for c in v))
    return [c / n for c in v]


def gaussian_bumps(seed, ngf, nbumps=8, amplitude=1e-3, wmin=1.0, wmax=3.0, cmin=-10.0, cmax=10.0):
    """[[(A, w, cx, cy, cz)] * nbumps] * ngf."""
    u = uniforms(seed, ngf * nbumps * 4)
    out = []
    k = 0
    for _ in range(ngf):
        row = []
        for _ in range(nbumps):
            w = wmin + (wmax - wmin) * u[k]
            c = [cmin + (cmax - cmin) * u[k + 1 + d] for d in range(3)]
            row.append((amplitude, w, c[0], c[1], c[2]))
            k += 4
        out.append(row)
    return out


def bssn_gaussian_state_numpy(xx, gf_names, seed):
    """State array (ngf, N2, N1, N0) on a Cartesian grid given xx = [x, y, z] arrays."""
    bumps = gaussian_bumps(seed, len(gf_names))
    X = xx[0][None, None, :]
    Y = xx[1][None, :, None]
    Z = xx[2][:, None, None]
    out = np.zeros((len(gf_names), len(xx[2]), len(xx[1]), len(xx[0])))
    for g, name in enumerate(gf_names):
        acc = np.full(out.shape[1:], 1.0 if name in BSSN_FLAT_ONE else 0.0)
        for A, w, cx, cy, cz in bumps[g]:
            acc = acc + A * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (w * w))
        out[g] = acc
    return out


def bssn_gaussian_state_torch(out, xx_dev, gf_names, seed):
    """Fill the device tensor ``out`` (ngf, N2, N1, N0) with the same state on the GPU."""
    bumps = gaussian_bumps(seed, len(gf_names))
    X = xx_dev[0][None, None, :]
    Y = xx_dev[1][None, :, None]
    Z = xx_dev[2][:, None, None]
    for g, name in enumerate(gf_names):
        acc = out[g]
        acc.fill_(1.0 if name in BSSN_FLAT_ONE else 0.0)
        for A, w, cx, cy, cz in bumps[g]:
            acc.add_(A * ((-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (w * w)).exp()))
    return out


def scalarwave_bump_parameters(seed, nbumps=16, cmin=-5.0, cmax=5.0, amin=0.5, amax=1.5):
    """Adjoint scalar-wave input: bump centres uniform in [cmin,cmax]^3 and amplitudes
    uniform in [amin,amax] (the amplitude range is this build's choice: the workload fixes
    only the centres and the width). Returns (amplitudes [n], centres [n][3])."""
    u = uniforms(seed, nbumps * 4)
    amps = [amin + (amax - amin) * u[4 * k] for k in range(nbumps)]
    cents = [[cmin + (cmax - cmin) * u[4 * k + 1 + d] for d in range(3)] for k in range(nbumps)]
    return amps, cents


def scalarwave_bump_basis_torch(xx_dev, centres, width=1.5):
    """(nbumps, N2, N1, N0) device tensor of exp(-|x - c_k|^2 / (2 width^2))."""
    import torch
    X = xx_dev[0][None, None, :]
    Y = xx_dev[1][None, :, None]
    Z = xx_dev[2][:, None, None]
    out = torch.empty((len(centres), len(xx_dev[2]), len(xx_dev[1]), len(xx_dev[0])), dtype=torch.float64,
                      device=xx_dev[0].device)
    for k, (cx, cy, cz) in enumerate(centres):
        out[k] = (-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2.0 * width * width)).exp()
    return out


def scalarwave_bump_basis_numpy(xx, centres, width=1.5):
    X = xx[0][None, None, :]
    Y = xx[1][None, :, None]
    Z = xx[2][:, None, None]
    return np.stack([np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2.0 * width * width))
                     for cx, cy, cz in centres])
