"""x-neighbours by cross-lane moves (codegen/lds.py xshift_kernel, nrpy_xshift in
nrpy_runtime_pre.h), on the CPU.

* The index logic, by brute force: a row of TX lanes (16, or the 64-lane wave), the first
  m of them live (m = TX, or 2 HX <= m < TX for a grid narrower than a row), lane l holding
  c = f(x0 + l) and h = f(x0 + l - HX) (l < HX) or f(x0 + l + HX). The moves are modelled as
  the hardware does them: lane l receives lane l + K of its own row (a chain of one-lane
  moves for the wave), 0 from a lane outside the row or not live. For every live lane and
  every offset 0 < |K| <= HX the selected value must be f(x0 + l + K). Controls: a halo
  taken from the wrong lane, and a row narrower than 2 HX, must fail.
* The pass on the generated Cartesian BSSN bundle: exactly the kernels named in the
  bundle's meta carry lane moves; in them no x-offset read of a moved gridfunction is left,
  every store goes through the ownership redirect, and the tile is moved to end at the
  grid's last point; the bundle rejects grids narrower than 2 HX.
"""
import json
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "nrpytutorial_amd", "csrc", "generated")


def _move(v, K, TX, m):
    """Lane l receives v of lane l + K of its row (0 if that lane is outside the row or
    not live); for TX = 64 a chain of |K| one-lane moves, each reading live lanes only."""
    n = len(v)
    out = v.copy()
    steps = [int(np.sign(K))] * abs(K) if TX == 64 else ([K] if K else [])
    for s in steps:
        nxt = np.zeros(n)
        for l in range(n):
            src = l + s
            same_row = 0 <= src < n and src // TX == l // TX
            nxt[l] = out[src] if same_row and (src % TX) < m else 0.0
        out = nxt
    return out


def _xshift(c, h, K, HX, TX, m, halo_shift=None):
    l = np.arange(len(c)) % TX
    in_row = _move(c, K, TX, m)
    hs = (K - HX if K > 0 else K + HX) if halo_shift is None else halo_shift(K, HX)
    halo = _move(h, hs, TX, m)
    keep = (l + K < m) if K > 0 else (l + K >= 0)
    return np.where(keep, in_row, halo)


def _row(TX, HX, m, nrows=2):
    """Values f(x) = distinct primes-ish per point, for rows of TX lanes at x0 = 100 r."""
    f = lambda x: 1.0 + 0.37 * x + 1e-3 * x * x
    c, h = np.zeros(nrows * TX), np.zeros(nrows * TX)
    x0 = np.repeat(100 * np.arange(nrows), TX)
    l = np.arange(nrows * TX) % TX
    live = l < m
    c[live] = f(x0 + l)[live]
    h[live] = np.where(l < HX, f(x0 + l - HX), f(x0 + l + HX))[live]
    return f, c, h, x0, l, live


@pytest.mark.parametrize("TX", [16, 64])
@pytest.mark.parametrize("HX", [2, 3])
def test_xshift_selects_the_neighbour_for_every_live_lane(TX, HX):
    for m in list(range(2 * HX, TX)) + [TX]:
        f, c, h, x0, l, live = _row(TX, HX, m, nrows=64 // TX if TX < 64 else 1)
        for K in [k for k in range(-HX, HX + 1) if k]:
            got = _xshift(c, h, K, HX, TX, m)
            want = f(x0 + l + K)
            assert np.array_equal(got[live], want[live]), (TX, HX, m, K)


def test_xshift_controls_fail():
    TX, HX = 16, 3
    f, c, h, x0, l, live = _row(TX, HX, TX)
    # halo from the wrong lane (one further along)
    wrong = lambda K, H: (K - H + 1 if K > 0 else K + H - 1)
    assert not np.array_equal(_xshift(c, h, 2, HX, TX, TX, wrong)[live], f(x0 + l + 2)[live])
    # a row narrower than 2 HX: the right halo is read from a lane that loaded the left halo
    m = 2 * HX - 1
    f, c, h, x0, l, live = _row(TX, HX, m)
    assert not np.array_equal(_xshift(c, h, 3, HX, TX, m)[live], f(x0 + l + 3)[live])


def _kernels(src):
    out = {}
    for m in re.finditer(r"__global__ void [^\n]*? (\w+)__gpu\(", src):
        start = m.end()
        nxt = src.find("__global__", start)
        out[m.group(1)] = src[start:nxt if nxt > 0 else len(src)]
    return out


def test_bundle_kernels_with_lane_moves():
    src = open(os.path.join(GEN, "bssn_cartesian.hip")).read()
    meta = json.load(open(os.path.join(GEN, "bssn_cartesian.json")))
    named = set(meta["xshift_kernels"])
    assert named, "no kernel uses lane moves"
    ks = _kernels(src)
    with_moves = {n for n, b in ks.items() if "nrpy_xshift<" in b}
    assert with_moves == named
    for n in named:
        body = ks[n]
        gfs = set(re.findall(r"const double nrpy_c\d+ = in_gfs\[IDX4\((\w+), i0,i1,i2\)\];", body))
        assert gfs
        for gf in gfs:      # no x-offset global read of a moved gridfunction is left
            assert not re.search(r"in_gfs\[IDX4\(%s, i0[+-]\d+,i1,i2\)\]" % gf, body), (n, gf)
        stores = [l for l in body.splitlines() if re.search(r"\]\s*=[^=]", l) and "nrpy_lds[" not in l
                  and not l.strip().startswith("const")]
        assert all("NRPY_OWNED" in l for l in stores), n
        assert "NRPY_RK4_STORE(" not in body
        assert "min(nrpy_xt, Nxx0 - " in body and "const bool nrpy_own" in body
    assert re.search(r"if \(p->Nxx0 < 6 \|\|", src)
