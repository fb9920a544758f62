"""Closed-form evolution of a plane wave under FD4 + the RK4 method of lines.

For u = 2 + sin(k.x - c t), v = du/dt, sampled on a uniform Cartesian grid,
the FD4 second derivative along axis d maps sin(k.x + phi) to itself times

    sigma_d = (2 (4/3) cos(k_d h_d) + 2 (-1/12) cos(2 k_d h_d) - 5/2) / h_d^2

(finite_difference.py:252-369 weights [-1/12, 4/3, -5/2, 4/3, -1/12]), so the
semi-discrete system for the oscillating part is (u, v)' = A (u, v) with
A = [[0, 1], [c^2 sum_d sigma_d, 0]], and one RK4 step multiplies it by the
classical amplification R = I + hA + (hA)^2/2 + (hA)^3/6 + (hA)^4/24
(MoLtimestepping/RK_Butcher_Table_Dictionary.py:86-91). The constant 2 is
annihilated by the stencil and stays. Points at least 8 cells (4 substeps x
stencil radius 2) away from every ghost zone do not see the boundary
conditions during one step, so there the grid values equal this closed form to
round-off. A wrong stage weight, a swapped time level in the diagonal schedule
(MoLtimestepping/C_Code_Generation.py:266-329), or a wrong stencil weight
changes R or sigma and fails the comparison.
"""
import numpy as np


def unit(k):
    k = np.asarray(k, float)
    return k / np.linalg.norm(k)


def amplification(dt, s, nsteps=1):
    A = np.array([[0.0, 1.0], [s, 0.0]])
    hA = dt * A
    R = np.eye(2) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    return np.linalg.matrix_power(R, nsteps)


def fd4_symbol(kd, h):
    th = kd * h
    return (2 * (4.0 / 3.0) * np.cos(th) + 2 * (-1.0 / 12.0) * np.cos(2 * th) - 2.5) / h ** 2


def plane_wave_after_rk4(xx, khat, dxx, c, dt, nsteps):
    """(u, v) arrays, shape (N2, N1, N0), on the coordinate arrays xx[0..2]."""
    s = c * c * sum(fd4_symbol(khat[d], dxx[d]) for d in range(3))
    R = amplification(dt, s, nsteps)
    X0, X1, X2 = xx[0][None, None, :], xx[1][None, :, None], xx[2][:, None, None]
    th = khat[0] * X0 + khat[1] * X1 + khat[2] * X2
    sn, cs = np.sin(th), np.cos(th)
    # initial (u, v) = (2 + sin, -c cos): sin-mode vector (1, 0), cos-mode vector (0, -c)
    u = 2 + R[0, 0] * sn + R[0, 1] * (-c) * cs
    v = R[1, 0] * sn + R[1, 1] * (-c) * cs
    return u, v


def pure_region(Nxx, G, nsteps):
    """Slices of points no boundary condition can reach within nsteps RK4 steps."""
    m = G + 8 * nsteps
    return tuple(slice(m, Nxx[d] + 2 * G - m) for d in (2, 1, 0))
