"""Vacuum Maxwell in rescaled curvilinear form (SinhCylindrical), the oracle against the
exact solution: the toroidal dipole (Maxwell/InitialData.py:25-113) is an exact solution
of the vacuum equations at every time, so the oracle's evolution
(VacuumMaxwell_Flat_Evol_Curvilinear_rescaled.py:11-196, RK4, the Sommerfeld branch of
Tutorial-Start_to_Finish-Solving_Maxwells_Equations_in_Vacuum-Curvilinear.ipynb with its
sinh_width = 0.4, CFL 0.5) must converge to it at fourth order; the reference checks the
same (72/96)^4-style scaling of E_x in that notebook (cell 46).

The dipole has only phi components (aU1, eU1); the others vanish identically and must
stay at rounding, and so must the divergence constraint C = Dhat_i E^i. The error is
measured at interior points within spherical radius 4 at t = 1, where the outgoing
pulse has not met the outer boundary.
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import Oracle, OracleGrid

PARS = dict(SINHWRHO=0.4, SINHWZ=0.4)


def _run(N, t_final):
    o = Oracle("maxwell_sinhcylindrical")
    Nxx = (N, 4, 2 * N)
    og = OracleGrid(o, Nxx, dict(PARS, time=0.0))
    y = og.zeros(8)
    og.initial_data(y)
    n = int(round(t_final / og.find_timestep(0.5)))
    og.mol_step_sommerfeld(y, t_final / n, n)
    ex_g = OracleGrid(o, Nxx, dict(PARS, time=t_final))
    ex = ex_g.zeros(8)
    ex_g.initial_data(ex)
    cart = np.zeros(3)
    R = np.zeros(og.shape)
    for i2 in range(og.shape[0]):
        for i0 in range(og.shape[2]):
            o.lib.oracle_xx_to_Cart(ctypes.byref(og.p), og.ptrs(og.xx), i0, 3, i2, cart.ctypes.data)
            R[i2, :, i0] = np.hypot(cart[0], cart[2])
    m = np.zeros(og.shape, bool)
    m[3:-3, 3:-3, 3:-3] = True
    m &= R < 4.0
    err = np.array([np.abs(y[g][m] - ex[g][m]).max() for g in range(8)])
    C = np.abs(og.maxwell_constraints(y)[0][m]).max()
    return err, C, o.evol


@pytest.fixture(scope="module")
def runs():
    return {N: _run(N, 1.0) for N in (24, 36, 48)}


def test_toroidal_dipole_converges_at_fourth_order(runs):
    evol = runs[24][2]
    phi = [evol.index("aU1"), evol.index("eU1")]
    for q in phi:
        e = [runs[N][0][q] for N in (24, 36, 48)]
        p1, p2 = np.log(e[0] / e[1]) / np.log(1.5), np.log(e[1] / e[2]) / np.log(48 / 36)
        print("%s: errors %s, orders %.2f %.2f" % (evol[q], e, p1, p2))
        assert 3.5 < p1 < 4.5 and 3.5 < p2 < 4.5
        assert e[2] < 1e-3
    for q in range(8):
        if q not in phi:
            assert max(runs[N][0][q] for N in (24, 36, 48)) < 1e-12, evol[q]
    assert max(runs[N][1] for N in (24, 36, 48)) < 1e-11
