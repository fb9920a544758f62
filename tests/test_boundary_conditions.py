"""Curvilinear boundary conditions on the CPU: boundary maps and ghost fill
against the reference's trusted SENR outputs, and the product's host-built
boundary lists against the oracle's, entry by entry.

Golden protocol (Tutorial-Start_to_Finish-Curvilinear_BCs.ipynb, cells 26-39):
SinhSphericalv2 / SinhCylindricalv2 with default parameters, FD order 4 so
NGHOSTS = 3, a 4x4x4 interior, ten AUX gridfunctions in alphabetical order
(rankoneU0..2, ranktwosymmDD00..22, rankzero: parity types 1..9 then 0, per
CurviBoundaryConditions.py:122-185), every value set to its linear index in
the gridfunction array, ghost zones overwritten with NaN, then one
apply_bcs_curvilinear; each line prints "i0 i1 i2 | " and the ten values cast
to int. The golden files are CurviBoundaryConditions/SENRout4x4x4-*.txt,
copied unchanged into tests/golden/curvibc/. SinhSpherical and SinhCylindrical
have the same EigenCoord (Spherical, Cylindrical) and the same 4x4x4 index
grid, so their boundary maps, and hence their outputs, must equal the same
golden files.
"""
import os

import numpy as np
import pytest

from oracle_lib import Oracle, OracleGrid
from nrpytutorial_amd import runtime as rt

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = {"scalarwave_sinhsphericalv2": "SENRout4x4x4-Spherical_NGHOSTS4oFD.txt",
          "scalarwave_sinhcylindricalv2": "SENRout4x4x4-Cylindrical_NGHOSTS4oFD.txt",
          "scalarwave_sinhspherical": "SENRout4x4x4-Spherical_NGHOSTS4oFD.txt",
          "scalarwave_sinhcylindrical": "SENRout4x4x4-Cylindrical_NGHOSTS4oFD.txt"}
PARITY = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9, 0], dtype=np.int8)
NGF = 10


def golden_text(bundle):
    with open(os.path.join(HERE, "golden", "curvibc", GOLDEN[bundle])) as f:
        return f.read()


def discrete_data(shape):
    """Linear index over the whole (gf, i2, i1, i0) array, ghosts set to NaN."""
    n2, n1, n0 = shape
    a = np.arange(NGF * n2 * n1 * n0, dtype=np.float64).reshape((NGF,) + shape)
    G = 3
    ghost = np.zeros(shape, bool)
    ghost[:G], ghost[-G:], ghost[:, :G], ghost[:, -G:], ghost[:, :, :G], ghost[:, :, -G:] = [True] * 6
    a[:, ghost] = np.nan
    return a


def print_like_playground(a):
    assert not np.isnan(a).any(), "NaN left in a ghost zone"
    _, n2, n1, n0 = a.shape
    lines = []
    for i2 in range(n2):
        for i1 in range(n1):
            for i0 in range(n0):
                lines.append("%d %d %d | " % (i0, i1, i2) + "".join("%d " % int(v) for v in a[:, i2, i1, i0]) + "\n")
    return "".join(lines)


@pytest.mark.parametrize("bundle", sorted(GOLDEN))
def test_oracle_bcs_match_senr_golden(bundle):
    o = Oracle(bundle)
    assert o.NGHOSTS == 3
    g = OracleGrid(o, (4, 4, 4), {})
    a = discrete_data(g.shape)
    g.apply_bcs(a, parity=PARITY.ctypes.data)
    assert print_like_playground(a) == golden_text(bundle)


def _host_lists(bundle, Nxx, **params):
    b = rt.Bundle(bundle)
    h = rt.HostBC(b, Nxx, **params)
    return b, h


@pytest.mark.parametrize("bundle", sorted(GOLDEN))
def test_product_composed_map_matches_senr_golden(bundle):
    """The one-pass ghost fill (each ghost point as a signed combination of
    interior points, per parity type) reproduces the golden exactly."""
    b, h = _host_lists(bundle, (4, 4, 4))
    gdest, rowptr, src, coeff, neg = h.composed()
    a = discrete_data((10, 10, 10))
    flat = a.reshape(NGF, -1)
    for gf in range(NGF):
        bit = 1 << int(PARITY[gf])
        c = np.where(neg & bit, -coeff, coeff)
        vals = np.add.reduceat(c * flat[gf, src], rowptr[:-1]) if len(src) else np.zeros(0)
        # rows with no terms do not occur: every ghost point has at least one source
        assert np.all(np.diff(rowptr) > 0)
        flat[gf, gdest] = vals
    assert print_like_playground(a) == golden_text(bundle)


def _compare_lists(h, og, G):
    ob, ib, nghost, _ = h.counts()
    oob, oib = og.counts()
    assert ob == oob and ib == oib
    for gz in range(G):
        outer, inner = h.layer(gz)
        o6, i16 = og.layer(gz)
        assert np.array_equal(np.array(outer, dtype=np.int32).reshape(-1, 6), o6)
        assert np.array_equal(np.array(inner, dtype=np.int32).reshape(-1, 16), i16)
    return nghost


@pytest.mark.parametrize("bundle,Nxx,params", [
    ("scalarwave_sinhsphericalv2", (4, 4, 4), {}),
    ("scalarwave_sinhcylindricalv2", (4, 4, 4), {}),
    ("scalarwave_sinhsphericalv2", (16, 8, 12), {}),
    ("scalarwave_sinhcylindricalv2", (12, 8, 10), {}),
    ("bssn_spherical_sym2", (72, 12, 2), {"RMAX": 7.5}),
    ("bssn_spherical_sym2", (24, 8, 2), {"RMAX": 7.5}),
    ("bssn_cartesian", (12, 10, 8), {}),
    ("scalarwave_cartesian", (6, 5, 4), {}),
    ("scalarwave_symtp", (12, 8, 10), {}),
    ("scalarwave_symtp", (16, 12, 4), {"AMAX": 3.0, "bScale": 0.8}),
    ("scalarwave_sinhspherical", (16, 8, 12), {}),
    ("scalarwave_sinhcylindrical", (12, 8, 10), {}),
    ("scalarwave_sinhsymtp", (12, 8, 10), {}),
    ("scalarwave_sinhsymtp", (16, 12, 4), {"AMAX": 3.0, "bScale": 0.8, "SINHWAA": 0.3}),
])
def test_product_bc_lists_equal_oracle(bundle, Nxx, params):
    """Boundary struct built by the product library on the host, entry by entry
    (destination, face direction, source, ten parities), equals the oracle's
    restatement of set_up__bc_gz_map_and_parity_condns.h + set_bcstruct.h."""
    b, h = _host_lists(bundle, Nxx, **params)
    og = OracleGrid(Oracle(bundle), Nxx, params)
    nghost = _compare_lists(h, og, b.NGHOSTS)
    n0, n1, n2 = (v + 2 * b.NGHOSTS for v in Nxx)
    assert nghost == n0 * n1 * n2 - Nxx[0] * Nxx[1] * Nxx[2]


@pytest.mark.parametrize("bundle,Nxx,params", [
    ("scalarwave_sinhsphericalv2", (16, 8, 12), {}),
    ("bssn_spherical_sym2", (72, 12, 2), {"RMAX": 7.5}),
    ("scalarwave_sinhcylindricalv2", (12, 8, 10), {}),
    ("scalarwave_symtp", (12, 8, 10), {}),
    ("scalarwave_sinhspherical", (16, 8, 12), {}),
    ("scalarwave_sinhcylindrical", (12, 8, 10), {}),
    ("scalarwave_sinhsymtp", (12, 8, 10), {}),
])
def test_composed_map_invariants(bundle, Nxx, params):
    """Every ghost point appears once, its sources are interior points, and for
    a scalar (parity type 0, all parities +1) the weights of each ghost point
    sum to 1: quadratic extrapolation (3, -3, 1) and parity copies preserve a
    constant field."""
    b, h = _host_lists(bundle, Nxx, **params)
    gdest, rowptr, src, coeff, neg = h.composed()
    G = b.NGHOSTS
    n0, n1, n2 = (v + 2 * G for v in Nxx)
    assert len(set(gdest.tolist())) == len(gdest)
    i0, i1, i2 = src % n0, (src // n0) % n1, src // (n0 * n1)
    assert np.all((i0 >= G) & (i0 < n0 - G) & (i1 >= G) & (i1 < n1 - G) & (i2 >= G) & (i2 < n2 - G))
    d0, d1, d2 = gdest % n0, (gdest // n0) % n1, gdest // (n0 * n1)
    inter = (d0 >= G) & (d0 < n0 - G) & (d1 >= G) & (d1 < n1 - G) & (d2 >= G) & (d2 < n2 - G)
    assert not inter.any()
    assert not (neg & 1).any()
    sums = np.add.reduceat(coeff, rowptr[:-1])
    assert np.allclose(sums, 1.0, rtol=0, atol=1e-12)
