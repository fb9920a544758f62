"""Adjoint path on the GPU: generated VJP kernels, transposed ghost fill, RK4
step adjoint and the checkpointed autograd evolution.

* Dot-product identities (exact for these linear systems up to rounding):
  <v, R x> = <R^T v, x> for rhs_eval_vjp (ScalarWave, Maxwell),
  <E x, y> = <x, E^T y> for bc_apply_T (Cartesian and SinhSphericalv2 maps,
  every parity type), <l, Phi d> = <Phi^T l, d> for MoL_step_adjoint.
  Tolerance 1e-12 relative to the magnitude of the summed terms.
* Adjoint scalar wave (128^3 on [-10,10]^3, 200 RK4 steps, CFL 0.5, loss
  J = 0.5 sum_interior uu(T)^2 dx^3, 16 Gaussian-bump amplitudes in the
  initial uu): gradient from torch autograd through RK4Evolve (sqrt(N)
  checkpointing) against central finite differences of the same GPU
  evolution, within 1e-6 relative. The evolution and the initial data are
  linear, so J is exactly quadratic in the amplitudes and the central
  difference has no truncation error for any h; its error is rounding in J
  divided by h. The test measures that rounding from the two central
  differences (2.8e-10 of J on MI355X; J at fixed inputs repeats bit for bit),
  so h=1e-6 gives only ~1e-2 agreement while h=1e-1 isolates the linear
  response; both are printed, h=1e-1 is asserted.
  Over 20 steps h=1e-6 already agrees to ~1e-8.
* The same gradient at 24^3 / 20 steps against central differences of the C
  oracle's evolution (CPU), within 1e-6 relative (h=1e-6 and h=1e-1).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(bundle, Nxx, **pars):
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    b = Bundle(bundle)
    g = Grid(b, Nxx, **pars)
    return b, g, Evolution(g)


def _rel(a, b, scale):
    return abs(a - b) / scale


@pytest.mark.parametrize("bundle", ["scalarwave_cartesian", "maxwell_cartesian"])
def test_rhs_vjp_dot_product(torch_cuda, bundle):
    torch = torch_cuda
    from nrpytutorial_amd.runtime import check, ptr
    print("case: rhs_eval_vjp dot product %s" % bundle, flush=True)
    b, g, ev = _grid(bundle, (14, 12, 10))
    n = len(b.evol)
    gen = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn((n,) + g.shape, generator=gen, dtype=torch.float64, device="cuda")
    v = torch.zeros_like(x)
    g.interior(v).copy_(g.interior(torch.randn(x.shape, generator=gen, dtype=torch.float64, device="cuda")))
    rx = g.empty(n)
    check(b.lib.rhs_eval(g.pp(), ptr(x), ptr(rx), ev.s()), "rhs_eval")
    G = g.empty(len(b.meta["vjp_slots"]))
    vt = g.empty(n)
    import ctypes
    f = b.lib.rhs_eval_vjp
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 8
    check(f(g.pp(), None, ptr(x), ptr(v), ptr(G), ptr(vt), None, ev.s()), "rhs_eval_vjp")
    torch.cuda.synchronize()
    lhs = float((g.interior(v) * g.interior(rx)).sum())
    rhs = float((vt * x).sum())
    scale = float((g.interior(v).abs() * g.interior(rx).abs()).sum())
    print("<v,Rx>=%.15e <R^T v,x>=%.15e" % (lhs, rhs))
    assert _rel(lhs, rhs, scale) < 1e-12


@pytest.mark.parametrize("bundle,Nxx", [("scalarwave_cartesian", (10, 9, 8)),
                                        ("scalarwave_sinhsphericalv2", (12, 8, 10)),
                                        ("scalarwave_sinhcylindricalv2", (10, 8, 12))])
def test_bc_transpose_dot_product(torch_cuda, bundle, Nxx):
    torch = torch_cuda
    import ctypes
    from nrpytutorial_amd.runtime import check, ptr
    print("case: bc_apply_T dot product %s" % bundle, flush=True)
    b, g, ev = _grid(bundle, Nxx)
    parity = (ctypes.c_int8 * 10)(1, 2, 3, 4, 5, 6, 7, 8, 9, 0)
    f = b.lib.bc_apply_T
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 6
    gen = torch.Generator(device="cuda").manual_seed(9)
    x = torch.randn((10,) + g.shape, generator=gen, dtype=torch.float64, device="cuda")
    y = torch.randn((10,) + g.shape, generator=gen, dtype=torch.float64, device="cuda")
    ex = x.clone()
    check(b.lib.apply_bcs_curvilinear_composed(g.pp(), g.bc, 10, parity, ptr(ex), ev.s()), "bc")
    ety = y.clone()
    check(f(g.pp(), g.bc, 10, parity, ptr(ety), ev.s()), "bc_apply_T")
    torch.cuda.synchronize()
    lhs = float((ex * y).sum())
    rhs = float((x * ety).sum())
    scale = float((ex.abs() * y.abs()).sum())
    G = b.NGHOSTS
    ghost = torch.ones(g.shape, dtype=torch.bool, device="cuda")
    ghost[G:-G, G:-G, G:-G] = False
    assert float(ety[:, ghost].abs().max()) == 0.0
    print("<Ex,y>=%.15e <x,E^T y>=%.15e" % (lhs, rhs))
    assert _rel(lhs, rhs, scale) < 1e-12


@pytest.mark.parametrize("bundle,pars", [("scalarwave_cartesian", {}),
                                         ("maxwell_cartesian", dict(xmin=-4.0, xmax=4.0, ymin=-4.0, ymax=4.0,
                                                                    zmin=-4.0, zmax=4.0))])
def test_step_adjoint_dot_product(torch_cuda, bundle, pars):
    torch = torch_cuda
    from nrpytutorial_amd.adjoint import RK4Adjoint
    print("case: MoL_step_adjoint dot product %s" % bundle, flush=True)
    b, g, ev = _grid(bundle, (16, 14, 12), **pars)
    adj = RK4Adjoint(ev)
    n = len(b.evol)
    gen = torch.Generator(device="cuda").manual_seed(13)
    d = torch.randn((n,) + g.shape, generator=gen, dtype=torch.float64, device="cuda")
    lam = torch.zeros_like(d)
    g.interior(lam).copy_(g.interior(torch.randn(d.shape, generator=gen, dtype=torch.float64, device="cuda")))
    dt = g.find_timestep(0.5)
    ev.y_n.copy_(d)
    ev.apply_bcs(ev.y_n, composed=True)
    d_bc = ev.y_n.clone()
    ev.step(dt, 1, use_graph=False)
    phid = ev.y_n.clone()
    l2 = lam.clone()
    adj.step_back(d_bc, l2, dt)
    adj.bc_T(l2)   # the initial ghost fill of d
    torch.cuda.synchronize()
    lhs = float((lam * phid).sum())
    rhs = float((l2 * d).sum())
    scale = float((lam.abs() * phid.abs()).sum())
    print("<l,Phi d>=%.15e <Phi^T l,d>=%.15e" % (lhs, rhs))
    assert _rel(lhs, rhs, scale) < 1e-12


def _scalarwave_loss_setup(torch, Nxx, nsteps, seed=3):
    from nrpytutorial_amd.adjoint import RK4Adjoint
    from nrpytutorial_amd.workloads import scalarwave_bump_parameters, scalarwave_bump_basis_torch
    b, g, ev = _grid("scalarwave_cartesian", Nxx, wavespeed=1.0)
    adj = RK4Adjoint(ev)
    amps, cents = scalarwave_bump_parameters(seed)
    basis = scalarwave_bump_basis_torch(g.xx_dev, cents)
    dt = 0.5 * min(g.params.dxx0, g.params.dxx1, g.params.dxx2)
    dV = g.params.dxx0 * g.params.dxx1 * g.params.dxx2
    return b, g, ev, adj, amps, cents, basis, dt, dV


def _loss(torch, g, ev, adj, basis, a, dt, nsteps, dV, grad):
    from nrpytutorial_amd.adjoint import RK4Evolve
    uu0 = 2.0 + (a[:, None, None, None] * basis).sum(0)
    y0 = torch.stack([uu0, torch.zeros_like(uu0)])
    yT = RK4Evolve().apply(y0, ev, adj, dt, nsteps)
    J = 0.5 * (g.interior(yT[0]) ** 2).sum() * dV
    if grad:
        J.backward()
    return J


def test_adjoint_scalarwave_128cubed_vs_central_differences(torch_cuda):
    torch = torch_cuda
    print("case: adjoint scalar wave 128^3, 200 steps, 16 amplitudes", flush=True)
    Nxx, nsteps = (128, 128, 128), 200
    b, g, ev, adj, amps, cents, basis, dt, dV = _scalarwave_loss_setup(torch, Nxx, nsteps)
    a = torch.tensor(amps, dtype=torch.float64, device="cuda", requires_grad=True)
    J = _loss(torch, g, ev, adj, basis, a, dt, nsteps, dV, grad=True)
    grad = a.grad.detach().cpu().numpy()
    fds, fdv = {}, {}
    with torch.no_grad():
        for h in (1e-6, 1e-1):
            fd = np.zeros(len(amps))
            for k in range(len(amps)):
                ap = torch.tensor(amps, dtype=torch.float64, device="cuda")
                am = ap.clone()
                ap[k] += h
                am[k] -= h
                Jp = float(_loss(torch, g, ev, adj, basis, ap, dt, nsteps, dV, grad=False))
                Jm = float(_loss(torch, g, ev, adj, basis, am, dt, nsteps, dV, grad=False))
                fd[k] = (Jp - Jm) / (2 * h)
            fds[h] = np.max(np.abs(grad - fd)) / np.max(np.abs(fd))
            fdv[h] = fd
            print("J=%.12e  max|grad|=%.6e  h=%g: rel err vs FD %.3e" % (float(J), np.max(np.abs(fd)), h, fds[h]))
        # J is exactly quadratic in the amplitudes, so both central differences are exact in h
        # and any difference between them is rounding in J: |Jp - Jm| error = |fd_h1 - fd_h2| * 2h.
        # Printed as a fraction of J, and the run-to-run spread of J at fixed inputs beside it.
        noise = np.max(np.abs(fdv[1e-6] - fdv[1e-1])) * 2e-6 / abs(float(J))
        J1 = float(_loss(torch, g, ev, adj, basis, a.detach(), dt, nsteps, dV, grad=False))
        J2 = float(_loss(torch, g, ev, adj, basis, a.detach(), dt, nsteps, dV, grad=False))
        print("rounding in J(a +- 1e-6 e_k) relative to J: %.3e; run-to-run |J1 - J2|/J at fixed a: %.3e"
              % (noise, abs(J1 - J2) / abs(J1)))
    assert fds[1e-1] < 1e-6


def test_adjoint_scalarwave_vs_oracle_central_differences(torch_cuda):
    """GPU adjoint gradient vs finite differences of the C oracle's evolution."""
    torch = torch_cuda
    from oracle_lib import Oracle, OracleGrid
    from nrpytutorial_amd.workloads import scalarwave_bump_basis_numpy
    print("case: adjoint scalar wave 24^3, 20 steps vs oracle FD", flush=True)
    Nxx, nsteps = (24, 24, 24), 20
    b, g, ev, adj, amps, cents, basis, dt, dV = _scalarwave_loss_setup(torch, Nxx, nsteps)
    a = torch.tensor(amps, dtype=torch.float64, device="cuda", requires_grad=True)
    _loss(torch, g, ev, adj, basis, a, dt, nsteps, dV, grad=True)
    grad = a.grad.detach().cpu().numpy()
    og = OracleGrid(Oracle("scalarwave_cartesian"), Nxx, dict(wavespeed=1.0))
    B = scalarwave_bump_basis_numpy(og.xx, cents)

    def J_oracle(av):
        y = og.zeros(2)
        y[0] = 2.0 + np.tensordot(av, B, axes=1)
        og.apply_bcs(y)
        og.mol_step(y, dt, nsteps)
        return 0.5 * np.sum(og.interior(y[0]) ** 2) * dV

    ks = [0, 5, 11, 15]
    for h in (1e-6, 1e-1):
        fd = []
        for k in ks:
            ap = np.array(amps)
            am = ap.copy()
            ap[k] += h
            am[k] -= h
            fd.append((J_oracle(ap) - J_oracle(am)) / (2 * h))
        fd = np.array(fd)
        rel = np.max(np.abs(grad[ks] - fd)) / np.max(np.abs(fd))
        print("h=%g: rel err vs oracle FD %.3e" % (h, rel))
        assert rel < 1e-6
