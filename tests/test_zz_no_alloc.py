"""gridfunction_sumsq_noalloc (the entry point that allocates nothing) against
gridfunction_sumsq (caller workspace), through the C ABI: bit-identical sums on two
calls in a row, the second on a smaller grid, so the scratch array's contents left by
the first call must not leak into the second. Sizes cover the block-count cap (a grid
with more than 1024 blocks of points) and a single-block grid."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sums(torch, b, Nxx, seed):
    from nrpytutorial_amd.runtime import Grid, ptr, check, stream_handle
    g = Grid(b, Nxx, kk0=1.0, kk1=0.0, kk2=0.0, wavespeed=1.0, time=0.0)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    gfs = torch.randn((2,) + g.shape, dtype=torch.float64, device="cuda", generator=gen)
    work = torch.zeros(1024, dtype=torch.float64, device="cuda")
    res = torch.zeros(4, dtype=torch.float64, device="cuda")
    s = stream_handle(torch)
    for gf in range(2):
        check(b.lib.gridfunction_sumsq(g.pp(), ptr(gfs), gf, ptr(work), ctypes.c_void_p(res.data_ptr() + 8 * gf), s),
              "gridfunction_sumsq")
        check(b.lib.gridfunction_sumsq_noalloc(g.pp(), ptr(gfs), gf, ctypes.c_void_p(res.data_ptr() + 8 * (2 + gf)),
                                               s), "gridfunction_sumsq_noalloc")
    torch.cuda.synchronize()
    r = res.cpu().numpy()
    want = [float((g.interior(gfs[q]) ** 2).sum()) for q in range(2)]
    return r[:2], r[2:], want


def test_noalloc_sum_matches_workspace_sum(torch_cuda):
    torch = torch_cuda
    from nrpytutorial_amd.runtime import Bundle
    b = Bundle("scalarwave_cartesian")
    for k, Nxx in enumerate([(96, 80, 72), (6, 5, 4)]):  # 552960 points (> 1024 blocks), then 120
        print("case: gridfunction_sumsq_noalloc vs gridfunction_sumsq, call %d, Nxx=%s" % (k + 1, Nxx), flush=True)
        ws, na, want = _sums(torch, b, Nxx, seed=k)
        print("  workspace %s noalloc %s torch %s" % (ws, na, want), flush=True)
        assert np.array_equal(ws, na)
        assert np.allclose(na, want, rtol=1e-12)


def test_noalloc_rejects_bad_arguments(torch_cuda):
    from nrpytutorial_amd.runtime import Bundle, Grid
    b = Bundle("scalarwave_cartesian")
    g = Grid(b, (8, 8, 8), kk0=1.0, kk1=0.0, kk2=0.0, wavespeed=1.0, time=0.0)
    assert b.lib.gridfunction_sumsq_noalloc(g.pp(), None, 0, None, None) == -1
    x = torch_cuda.zeros(1, dtype=torch_cuda.float64, device="cuda")
    assert b.lib.gridfunction_sumsq_noalloc(g.pp(), ctypes.c_void_p(x.data_ptr()), -1,
                                            ctypes.c_void_p(x.data_ptr()), None) == -4
