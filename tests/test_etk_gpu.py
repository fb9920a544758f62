"""Einstein Toolkit plugin form of the BSSN kernels on the GPU, through the C ABI.

Each gridfunction is its own array of the toolkit's allocated shape cctk_ash,
padded beyond cctk_lsh with NaN, and every output array starts as NaN. On the
seeded flat + Gaussian state of the SENR-layout tests:

* Ricci tensor, BSSN right-hand sides, H and MU at the interior points
  [cctk_nghostzones, cctk_lsh - cctk_nghostzones) equal the SENR-layout
  kernels' values bit for bit (the kernels are the same arithmetic; see
  tests/test_etk.py), except the outputs of the SENR group that builds its
  x-neighbours by lane moves, whose reordered statements the compiler contracts
  into FMAs differently: those agree to 1e-13 of each gridfunction's max norm;
  and the right-hand sides agree with the C oracle to 1e-12;
* every output point outside that range is still NaN (nothing written there),
  and no output is NaN inside it (no padding or out-of-range point was read);
* det(gammabar) enforcement at every point equals the SENR kernel's bit for bit.

The cases cover allocated rows of even and odd length, ghost zones deeper than
the stencils need (so the interior no longer starts at NGHOSTS), an odd
cctk_lsh[0] inside an even cctk_ash[0] (the last point of a row must still
reach LDS), and partial 16x4x4 and 64x1x4 tiles: both the direct global-to-LDS copy and
the copy through registers run.
"""
import numpy as np
import pytest

from helpers import rel_maxnorm_err, closer_to_exact

pytestmark = pytest.mark.gpu
TOL = 1e-12

# (SENR interior Nxx, cctk_nghostzones, padding ash - lsh)
CASES = [((14, 12, 10), (3, 3, 3), (0, 0, 0)),
         ((13, 9, 7), (3, 3, 3), (2, 1, 0)),
         ((14, 12, 10), (4, 3, 5), (2, 3, 1)),
         ((16, 12, 10), (5, 3, 3), (0, 2, 0)),
         ((13, 9, 7), (3, 3, 3), (1, 0, 0))]


def _setup(torch, Nxx, ng, pad):
    from nrpytutorial_amd.runtime import Bundle, Grid, Evolution
    from nrpytutorial_amd.workloads import bssn_gaussian_state_numpy
    from nrpytutorial_amd.etk import Baikal, cctk_grid
    b = Bundle("bssn_cartesian")
    g = Grid(b, Nxx)
    ev = Evolution(g)
    y = bssn_gaussian_state_numpy(g.xx, b.evol, 7)
    ev.y_n.copy_(torch.from_numpy(y))
    k = Baikal()
    lsh = tuple(n + 2 * b.NGHOSTS for n in Nxx)
    ash = tuple(l + p for l, p in zip(lsh, pad))
    gh = cctk_grid(lsh, ng, (g.params.dxx0, g.params.dxx1, g.params.dxx2), ash)
    params = k.new_params(eta=g.params.eta)

    def arrays(n, src=None):
        out = []
        for q in range(n):
            a = torch.full((ash[2], ash[1], ash[0]), float("nan"), dtype=torch.float64, device="cuda")
            if src is not None:
                a[:lsh[2], :lsh[1], :lsh[0]] = src[q]
            out.append(a)
        return out
    return b, g, ev, y, k, gh, params, lsh, arrays


def _split(arrs, lsh, ng):
    """(values at the interior points, values at every other point of [0, lsh))."""
    inner = np.zeros((lsh[2], lsh[1], lsh[0]), bool)
    inner[ng[2]:lsh[2] - ng[2], ng[1]:lsh[1] - ng[1], ng[0]:lsh[0] - ng[0]] = True
    vals = np.stack([a.cpu().numpy()[:lsh[2], :lsh[1], :lsh[0]] for a in arrs])
    return vals[:, inner], vals[:, ~inner], inner


@pytest.mark.parametrize("Nxx,ng,pad", CASES)
def test_plugin_form_equals_senr_kernels(torch_cuda, Nxx, ng, pad):
    torch = torch_cuda
    from nrpytutorial_amd.runtime import ptr, check
    print("case: bssn_etk Nxx=%s nghostzones=%s padding=%s" % (Nxx, ng, pad), flush=True)
    b, g, ev, y, k, gh, params, lsh, arrays = _setup(torch, Nxx, ng, pad)
    # SENR-layout reference values
    check(b.lib.Ricci_eval(g.rfm, g.pp(), ptr(ev.y_n), ptr(ev.auxevol), ev.s()), "Ricci_eval")
    rhs = g.empty(len(b.evol))
    check(b.lib.rhs_eval(g.rfm, g.pp(), ptr(ev.auxevol), ptr(ev.y_n), ptr(rhs), ev.s()), "rhs_eval")
    H = ev.hamiltonian()
    MU = ev.momentum()
    torch.cuda.synchronize()
    # plugin form
    evol = arrays(len(k.evol), ev.y_n)
    rbar = arrays(len(k.auxevol))
    rhs_e = arrays(len(k.evol))
    aux = arrays(len(k.aux))
    k.ricci(gh, params, evol, rbar)
    k.rhss(gh, params, evol, rbar, rhs_e)
    k.constraints(gh, params, evol, aux)
    torch.cuda.synchronize()
    # The SENR kernels of an output group that builds its x-neighbours by lane moves
    # (bundle meta "xshift_kernels") order their statements differently, so the compiler
    # contracts other multiply-adds into FMAs there: those outputs agree to rounding
    # (1e-13 of each gridfunction's max norm), every other output bit for bit.
    moved = [n for k, grp in enumerate(b.meta["rhs_groups"])
             if "rhs_eval_g%d" % k in b.meta.get("xshift_kernels", []) for n in grp]
    for name, got, want in (("RbarDD", rbar, ev.auxevol), ("rhs", rhs_e, rhs),
                            ("H, MU", aux, torch.cat([H[None], MU]))):
        inside, outside, inner = _split(got, lsh, ng)
        ref = want.cpu().numpy()[:, inner]
        assert not np.isnan(inside).any(), name
        assert np.isnan(outside).all(), name
        if name != "rhs":
            assert np.array_equal(inside, ref), (name, np.abs(inside - ref).max())
            continue
        for q, gf in enumerate(b.evol):
            if gf in moved:
                e = np.abs(inside[q] - ref[q]).max() / max(np.abs(ref[q]).max(), 1e-300)
                assert e <= 1e-13, (gf, e)
            else:
                assert np.array_equal(inside[q], ref[q]), (gf, np.abs(inside[q] - ref[q]).max())
    # right-hand sides against the oracle (its Ricci tensor, its RHS): 1e-12 of each
    # gridfunction's max norm, or, where the oracle's own FP64 rounding is that large (second
    # derivatives of the ~1 lapse and conformal factor, which the plugin kernels difference
    # from the centre value), within 2x the oracle's distance to the extended-precision
    # oracle's value
    from oracle_lib import Oracle, OracleGrid
    og = OracleGrid(Oracle("bssn_cartesian"), Nxx, {})
    ric_ref = og.ricci(y)
    rhs_ref = og.rhs_eval(y, ric_ref)
    inside, _, inner = _split(rhs_e, lsh, ng)
    err = rel_maxnorm_err(inside, rhs_ref[:, inner])
    print("rhs vs oracle %.2e" % err)
    if err >= TOL:
        ogx = OracleGrid(Oracle("bssn_cartesian_ld"), Nxx, {})
        exact = ogx.rhs_eval(y, ogx.ricci(y))[:, inner]
        ok, e_got, e_ref = closer_to_exact(inside, rhs_ref[:, inner], exact)
        print("vs extended precision: plugin %.2e, oracle %.2e" % (e_got.max(), e_ref.max()))
        assert ok, (e_got, e_ref)
    # the inputs are untouched, padding included
    for q, a in enumerate(evol):
        v = a.cpu().numpy()
        assert np.array_equal(v[:lsh[2], :lsh[1], :lsh[0]], y[q])
        v[:lsh[2], :lsh[1], :lsh[0]] = np.nan
        assert np.isnan(v).all()


@pytest.mark.parametrize("Nxx,ng,pad", CASES[:2])
def test_plugin_detgammabar_equals_senr_kernel(torch_cuda, Nxx, ng, pad):
    torch = torch_cuda
    from nrpytutorial_amd.runtime import ptr, check
    print("case: bssn_etk detgammabar Nxx=%s padding=%s" % (Nxx, pad), flush=True)
    b, g, ev, y, k, gh, params, lsh, arrays = _setup(torch, Nxx, ng, pad)
    evol = arrays(len(k.evol), ev.y_n)
    check(b.lib.enforce_detgammabar_constraint(g.rfm, g.pp(), ptr(ev.y_n), ev.s()), "enforce_detgammabar_constraint")
    k.enforce_detgammabar(gh, params, evol)
    torch.cuda.synchronize()
    want = ev.y_n.cpu().numpy()
    assert not np.array_equal(want, y)
    for q, a in enumerate(evol):
        v = a.cpu().numpy()
        assert np.array_equal(v[:lsh[2], :lsh[1], :lsh[0]], want[q]), k.evol[q]
        v[:lsh[2], :lsh[1], :lsh[0]] = np.nan
        assert np.isnan(v).all()
