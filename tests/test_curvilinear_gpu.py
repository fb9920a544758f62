"""Curvilinear scalar wave and boundary conditions on the GPU, and the RK4
closed form on the plain scalar-wave path, through the bundles' C ABI.

* Boundary conditions: the device ghost fills (sequential per-layer kernels
  and the one-launch composed map) reproduce the reference's trusted SENR
  outputs exactly (CurviBoundaryConditions/SENRout4x4x4-*.txt protocol, see
  tests/test_boundary_conditions.py).
* Scalar wave in SinhSpherical(v2), SinhCylindrical(v2), SymTP and SinhSymTP
  coordinates: initial data, rhs_eval with precomputed rfm tables, and an RK4
  evolution agree with the oracle (closed-form reference metric at every
  point) to 1e-12 / 1e-10.
* The plain 64^3 scalar wave (and a 192^3 grid) after RK4 steps equals the
  closed-form FD4 + RK4 amplification of a plane wave at every point the
  boundary conditions cannot reach.
"""
import numpy as np
import pytest

from helpers import rel_maxnorm_err, unit_direction
import mol_closed_form as cf
from test_boundary_conditions import GOLDEN, PARITY, discrete_data, golden_text, print_like_playground

pytestmark = pytest.mark.gpu
TOL = 1e-12
TOL_EVOL = 1e-10
TOL_CURVI_RHS = 1e-10


@pytest.mark.parametrize("composed", [False, True])
@pytest.mark.parametrize("bundle", sorted(GOLDEN))
def test_device_bcs_match_senr_golden(torch_cuda, bundle, composed):
    torch = torch_cuda
    import ctypes
    from nrpytutorial_amd.runtime import Bundle, Grid, check, ptr
    print("case: SENR golden %s composed=%s" % (bundle, composed), flush=True)
    b = Bundle(bundle)
    g = Grid(b, (4, 4, 4))
    a = discrete_data(g.shape)
    t = torch.from_numpy(a).cuda()
    par = (ctypes.c_int8 * 10)(*PARITY.tolist())
    f = b.lib.apply_bcs_curvilinear_composed if composed else b.lib.apply_bcs_curvilinear
    check(f(g.pp(), g.bc, 10, par, ptr(t), None), "apply_bcs")
    torch.cuda.synchronize()
    assert print_like_playground(t.cpu().numpy()) == golden_text(bundle)


# The sinh maps run with milder widths than their defaults (SINHW = 0.2 puts the first
# radial point at r ~ 0.02, where the 1/r^2 cancellation below amplifies rounding past
# TOL_CURVI_RHS); the SinhSphericalv2 / SinhCylindricalv2 bundles keep their defaults.
CURVI_CASES = [("scalarwave_sinhsphericalv2", (16, 8, 16), {}),
               ("scalarwave_sinhcylindricalv2", (12, 10, 14), {}),
               ("scalarwave_symtp", (16, 8, 12), {}),
               ("scalarwave_sinhspherical", (16, 8, 16), dict(AMPL=3.0, SINHW=0.4)),
               ("scalarwave_sinhcylindrical", (12, 10, 14), dict(AMPLRHO=3.0, SINHWRHO=0.4, AMPLZ=3.0, SINHWZ=0.4)),
               ("scalarwave_sinhsymtp", (16, 8, 12), dict(AMAX=3.0, SINHWAA=0.4))]


def _curvi(bundle, Nxx, seed=1, **extra):
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    from oracle_lib import Oracle, OracleGrid
    k = unit_direction(seed)
    pars = dict(kk0=k[0], kk1=k[1], kk2=k[2], wavespeed=1.0, time=0.0, **extra)
    b = Bundle(bundle)
    g = Grid(b, Nxx, **pars)
    return b, g, Evolution(g), OracleGrid(Oracle(bundle), Nxx, pars)


@pytest.mark.parametrize("bundle,Nxx,extra", CURVI_CASES, ids=[c[0] for c in CURVI_CASES])
def test_curvilinear_initial_data_and_rhs(torch_cuda, bundle, Nxx, extra):
    torch = torch_cuda
    from nrpytutorial_amd.runtime import check, ptr
    print("case: %s initial data + rhs_eval Nxx=%s" % (bundle, Nxx), flush=True)
    b, g, ev, og = _curvi(bundle, Nxx, **extra)
    ev.initial_data(apply_bcs=False)
    torch.cuda.synchronize()
    y_ref = og.zeros(2)
    og.initial_data(y_ref)
    assert rel_maxnorm_err(ev.y_n.cpu().numpy(), y_ref) < TOL
    rhs = g.empty(2)
    check(b.lib.rhs_eval(g.rfm, g.pp(), ptr(ev.y_n), ptr(rhs), ev.s()), "rhs_eval")
    torch.cuda.synchronize()
    rhs_ref = og.interior(og.rhs_eval(y_ref))
    got = g.interior(rhs.cpu().numpy())
    err = rel_maxnorm_err(got, rhs_ref)
    # Outer radial half: well-conditioned terms, reassociation-level agreement.
    n0 = Nxx[0]
    err_outer = rel_maxnorm_err(got[..., n0 // 2:], rhs_ref[..., n0 // 2:])
    print("rhs rel err: all %.3e, outer radial half %.3e" % (err, err_outer), flush=True)
    assert err_outer < TOL
    # Near the origin and the axis the curvilinear Laplacian sums terms with 1/r^2 and
    # cot(theta)/r^2 factors that cancel; there the device (precomputed tables, fused
    # multiply-adds) and the oracle (closed form per point, separate multiply and add)
    # differ by rounding amplified by that cancellation.
    assert err < TOL_CURVI_RHS


@pytest.mark.parametrize("bundle,Nxx,extra", CURVI_CASES, ids=[c[0] for c in CURVI_CASES])
def test_curvilinear_rk4_matches_oracle(torch_cuda, bundle, Nxx, extra):
    torch = torch_cuda
    print("case: %s RK4 8 steps" % bundle, flush=True)
    b, g, ev, og = _curvi(bundle, Nxx, **extra)
    ev.initial_data()
    y_ref = og.zeros(2)
    og.initial_data(y_ref)
    og.apply_bcs(y_ref)
    dt = g.find_timestep(0.5)
    assert abs(dt - og.find_timestep(0.5)) <= 1e-15 * dt
    ev.step(dt, nsteps=8, use_graph=True)
    og.mol_step(y_ref, dt, nsteps=8)
    torch.cuda.synchronize()
    err = rel_maxnorm_err(ev.y_n.cpu().numpy(), y_ref)
    print("rel err after 8 steps: %.3e" % err)
    assert err < TOL_EVOL


@pytest.mark.parametrize("Nxx,nsteps", [((64, 64, 64), 1), ((64, 64, 64), 2), ((192, 192, 192), 1)])
def test_scalarwave_rk4_closed_form(torch_cuda, Nxx, nsteps):
    """Size-independent check of the HIP RK4 path at the plain scalar-wave size
    (64^3, Cartesian [-10,10]^3, dt = 0.5 dx) and at 192^3."""
    torch = torch_cuda
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    print("case: scalarwave closed-form RK4 Nxx=%s nsteps=%d" % (Nxx, nsteps), flush=True)
    k = unit_direction(0)
    b = Bundle("scalarwave_cartesian")
    g = Grid(b, Nxx, kk0=k[0], kk1=k[1], kk2=k[2], wavespeed=1.0, time=0.0)
    ev = Evolution(g)
    ev.initial_data()
    dxx = (g.params.dxx0, g.params.dxx1, g.params.dxx2)
    dt = 0.5 * dxx[0]
    ev.step(dt, nsteps=nsteps, use_graph=True)
    torch.cuda.synchronize()
    y = ev.y_n.cpu().numpy()
    u, v = cf.plane_wave_after_rk4(g.xx, cf.unit(k), dxx, 1.0, dt, nsteps)
    sl = cf.pure_region(Nxx, b.NGHOSTS, nsteps)
    assert np.max(np.abs(y[0][sl] - u[sl])) < 1e-13
    assert np.max(np.abs(y[1][sl] - v[sl])) < 1e-13
