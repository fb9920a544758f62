"""The C oracle's method-of-lines evolution against closed forms (CPU).

* RK4 amplification: one and three RK4 steps of a plane wave equal the closed
  form of tests/mol_closed_form.py at every point the boundary conditions
  cannot reach.
* Convergence to the exact plane wave u = 2 + sin(k.x - ct) at fourth order
  (the reference's Start-to-Finish ScalarWave validation,
  Tutorial-Start_to_Finish-ScalarWave.ipynb and
  Tutorial-Start_to_Finish-ScalarWaveCurvilinear.ipynb: 16 -> 24 points and
  the (16/24)^4 rescaling), in Cartesian and SinhSphericalv2 coordinates.
"""
import numpy as np
import pytest

from oracle_lib import Oracle, OracleGrid
import mol_closed_form as cf

KVEC = (0.37, -0.81, 0.46)


def _plane_wave_grid(bundle, Nxx, **kw):
    o = Oracle(bundle)
    khat = cf.unit(KVEC)
    p = dict(kk0=khat[0], kk1=khat[1], kk2=khat[2], wavespeed=1.0, **kw)
    g = OracleGrid(o, Nxx, p)
    y = g.zeros(2)
    g.initial_data(y)
    g.apply_bcs(y)
    return o, g, y, khat


@pytest.mark.parametrize("nsteps", [1, 3])
def test_rk4_amplification_closed_form(nsteps):
    Nxx = (16 * nsteps + 20, 16 * nsteps + 18, 16 * nsteps + 22)
    o, g, y, khat = _plane_wave_grid("scalarwave_cartesian", Nxx)
    dxx = (g.p.dxx0, g.p.dxx1, g.p.dxx2)
    dt = 0.5 * min(dxx)
    g.mol_step(y, dt, nsteps)
    u, v = cf.plane_wave_after_rk4(g.xx, khat, dxx, 1.0, dt, nsteps)
    sl = cf.pure_region(Nxx, o.NGHOSTS, nsteps)
    assert np.max(np.abs(y[0][sl] - u[sl])) < 1e-13
    assert np.max(np.abs(y[1][sl] - v[sl])) < 1e-13
    # the closed form itself differs from the exact solution by the truncation error
    t = nsteps * dt
    X0, X1, X2 = g.xx[0][None, None, :], g.xx[1][None, :, None], g.xx[2][:, None, None]
    exact = 2 + np.sin(khat[0] * X0 + khat[1] * X1 + khat[2] * X2 - t)
    assert np.max(np.abs(y[0][sl] - exact[sl])) > 1e-9


def _max_err(bundle, Nxx, T, CFL, region, **kw):
    """Max |u - exact| at time T over ``region`` of the interior (a slice tuple in
    (i2, i1, i0) order): the extrapolated outer boundary of an outgoing wave is a
    low-order error source, so convergence is measured where it has not arrived."""
    o, g, y, khat = _plane_wave_grid(bundle, Nxx, **kw)
    dt = g.find_timestep(CFL)
    n = int(round(T / dt))
    dt = T / n
    g.mol_step(y, dt, n)
    g2 = OracleGrid(o, Nxx, dict(kk0=khat[0], kk1=khat[1], kk2=khat[2], wavespeed=1.0, time=T, **kw))
    ex = g2.zeros(2)
    g2.initial_data(ex)
    return np.max(np.abs(g.interior(y[0])[region] - g.interior(ex[0])[region]))


CART = dict(xmin=-3.0, xmax=3.0, ymin=-3.0, ymax=3.0, zmin=-3.0, zmax=3.0)


def test_cartesian_fourth_order_convergence():
    """Central quarter (per axis) of [-3,3]^3 after t=1; 16 points per wavelength at N=16."""
    errs = []
    for n in (16, 24, 32):
        q = n // 8
        errs.append(_max_err("scalarwave_cartesian", (n, n, n), 1.0, 0.5, (slice(n // 2 - q, n // 2 + q),) * 3,
                             **CART))
    r1 = np.log(errs[0] / errs[1]) / np.log(24 / 16)
    r2 = np.log(errs[1] / errs[2]) / np.log(32 / 24)
    assert 3.4 < r1 < 4.6 and 3.4 < r2 < 4.6, (errs, r1, r2)


def test_sinhsphericalv2_fourth_order_convergence():
    """Inner half in the radial index (all angles: the parity boundary conditions
    at the axis and origin are in the region) after t=0.5, AMPL=3."""
    errs = []
    for n in (8, 12, 16):
        errs.append(_max_err("scalarwave_sinhsphericalv2", (2 * n, n, 2 * n), 0.5, 1.0,
                             (slice(None), slice(None), slice(0, n)), AMPL=3.0))
    r1 = np.log(errs[0] / errs[1]) / np.log(12 / 8)
    r2 = np.log(errs[1] / errs[2]) / np.log(16 / 12)
    assert 3.4 < r1 < 4.6 and 3.4 < r2 < 4.6, (errs, r1, r2)


MAXWELL_BOX = dict(xmin=-4.0, xmax=4.0, ymin=-4.0, ymax=4.0, zmin=-4.0, zmax=4.0)


def test_maxwell_fourth_order_convergence():
    """Vacuum Maxwell System II, toroidal dipole (an exact solution for all t,
    Maxwell/InitialData.py), Cartesian [-4,4]^3, evolved to t=0.5 with CFL 0.5 at
    24^3, 32^3, 48^3: the state error and the constraint div E both fall at
    fourth order."""
    o = Oracle("maxwell_cartesian")
    errs, cons = [], []
    for n in (24, 32, 48):
        g = OracleGrid(o, (n, n, n), MAXWELL_BOX)
        y = g.zeros(8)
        g.initial_data(y)
        g.apply_bcs(y)
        T = 0.5
        dt = g.find_timestep(0.5)
        nst = int(round(T / dt))
        g.mol_step(y, T / nst, nst)
        g2 = OracleGrid(o, (n, n, n), dict(MAXWELL_BOX, time=T))
        ex = g2.zeros(8)
        g2.initial_data(ex)
        errs.append(np.max(np.abs(g.interior(y) - g.interior(ex))))
        cons.append(np.max(np.abs(g.interior(g.maxwell_constraints(y)[0]))))
    # div E involves one more derivative of the error, so its coarse-pair rate is lower.
    for e, lo in ((errs, 3.4), (cons, 3.2)):
        r1 = np.log(e[0] / e[1]) / np.log(32 / 24)
        r2 = np.log(e[1] / e[2]) / np.log(48 / 32)
        assert lo < r1 < 4.6 and lo < r2 < 4.6, (e, r1, r2)


def test_symtp_fourth_order_convergence():
    """SymTP (prolate spheroidal, foci at z = +-bScale; 2-D rfm tables in the bundle),
    AMAX=3: inner half in xx0 after t=0.5, all angles (parity boundaries at the axis
    and at A=0 are in the region)."""
    errs = []
    for n in (8, 12, 16):
        errs.append(_max_err("scalarwave_symtp", (2 * n, n, 2 * n), 0.5, 1.0,
                             (slice(None), slice(None), slice(0, n)), AMAX=3.0))
    r1 = np.log(errs[0] / errs[1]) / np.log(12 / 8)
    r2 = np.log(errs[1] / errs[2]) / np.log(16 / 12)
    print(errs, r1, r2)
    assert 3.4 < r1 < 4.6 and 3.4 < r2 < 4.6, (errs, r1, r2)


SINH_CASES = [
    ("scalarwave_sinhspherical", (8, 12, 16), dict(AMPL=3.0, SINHW=0.4)),
    ("scalarwave_sinhcylindrical", (12, 16, 24), dict(AMPLRHO=3.0, SINHWRHO=0.4, AMPLZ=3.0, SINHWZ=0.4)),
    ("scalarwave_sinhsymtp", (8, 12, 16), dict(AMAX=3.0, SINHWAA=0.4)),
]


@pytest.mark.parametrize("bundle,ns,kw", SINH_CASES, ids=[c[0] for c in SINH_CASES])
def test_sinh_maps_fourth_order_convergence(bundle, ns, kw):
    """SinhSpherical, SinhCylindrical and SinhSymTP (reference_metric.py sinh radial maps):
    plane wave to t=0.5 at CFL 1. Spherical-like systems: inner half in xx0, all angles.
    SinhCylindrical: inner half in rho and the middle half in z (the z faces are outer
    boundaries)."""
    errs = []
    for n in ns:
        if "cylindrical" in bundle:
            Nxx, reg = (n, n, 2 * n), (slice(n // 2, 3 * n // 2), slice(None), slice(0, n // 2))
        else:
            Nxx, reg = (2 * n, n, 2 * n), (slice(None), slice(None), slice(0, n))
        errs.append(_max_err(bundle, Nxx, 0.5, 1.0, reg, **kw))
    r1 = np.log(errs[0] / errs[1]) / np.log(ns[1] / ns[0])
    r2 = np.log(errs[1] / errs[2]) / np.log(ns[2] / ns[1])
    print(errs, r1, r2)
    assert 3.4 < r1 < 4.6 and 3.4 < r2 < 4.6, (errs, r1, r2)
