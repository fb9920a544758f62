"""Einstein Toolkit plugin form of the BSSN kernels, on the CPU.

* ``gfaccess`` in ETK mode spells accesses exactly as grid.py:92-100 does
  (``hDD00GF[CCTK_GFINDEX3D(cctkGH, i0, i1, i2)]``, ``_rhsGF`` for rhs_gfs) and
  refuses a non-3-D grid.
* The plugin-form kernels compute what the SENR-layout kernels compute: after
  the memory accesses of both are reduced to (gridfunction, RHS or not,
  offset), every statement after the bounds check is identical, LDS staging
  included (as a multiset for a SENR kernel that builds its x-neighbours by lane
  moves, which also reorders its stencil-level statements; the lines that switch a
  kernel to tile-relative addressing are left out). So the GPU parity of the SENR kernels against the oracle carries
  over, and the GPU tests check the values as well.
* The entry points reject bad grids, NULL pointers and FD orders other than 4
  on the host, before anything reaches the GPU.
"""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "nrpytutorial_amd", "csrc", "generated")

GFACCESS = """
import nrpytutorial_amd.NRPy_param_funcs as par
import nrpytutorial_amd.grid as gri
gri.register_gridfunctions("EVOL", ["uu"])
gri.register_gridfunctions("AUX", ["H"])
par.set_parval_from_str("grid::GridFuncMemAccess", "ETK")
print(gri.gfaccess("in_gfs", "uu"))
print(gri.gfaccess("rhs_gfs", "uu"))
print(gri.gfaccess("aux_gfs", "H", "i0+1,i1,i2-2"))
par.set_parval_from_str("grid::GridFuncMemAccess", "SENRlike")
print(gri.gfaccess("rhs_gfs", "uu"))
par.set_parval_from_str("grid::GridFuncMemAccess", "ETK")
par.set_parval_from_str("grid::DIM", 2)
try:
    gri.gfaccess("in_gfs", "uu")
except ValueError as e:
    print("ValueError")
"""


def test_gfaccess_etk_spelling():
    out = subprocess.run([sys.executable, "-c", GFACCESS], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.splitlines() == [
        "uuGF[CCTK_GFINDEX3D(cctkGH, i0, i1, i2)]",
        "uu_rhsGF[CCTK_GFINDEX3D(cctkGH, i0, i1, i2)]",
        "HGF[CCTK_GFINDEX3D(cctkGH, i0+1,i1,i2-2)]",
        "rhs_gfs[IDX4(UUGF, i0, i1, i2)]",
        "ValueError"]


def _kernel(src, name):
    start = src.index("%s__gpu(" % name)
    end = src.index("\n}\n", start)
    return src[start:end]


_SENR = re.compile(r"\b(\w+)\[IDX4\((\w+)GF, ([^\]]*?)\)\]")
_ETK = re.compile(r"\b(\w+?)(_rhs)?GF\[CCTK_GFINDEX3D\(cctkGH, ([^\]]*?)\)\]")


_XC = re.compile(r"const double nrpy_c(\d+) = in_gfs\[IDX4\((\w+)GF, i0,i1,i2\)\];")
_XSHIFT = re.compile(r"nrpy_xshift<\d+, \d+, (-?\d+)>\(nrpy_c(\d+), nrpy_h\d+, nrpy_xl, nrpy_xm\)")


def _undo_xshift(body):
    """A SENR kernel whose x-neighbours are built by lane moves, with every lane-move value
    written back as the read it replaces: ACC(GF, 0, i0+k,i1,i2). Its centre and halo loads
    and the ownership redirect of its stores are dropped."""
    gf = {q: g for q, g in _XC.findall(body)}
    body = _XSHIFT.sub(lambda m: "in_gfs[IDX4(%sGF, i0%+d,i1,i2)]" % (gf[m.group(2)], int(m.group(1))), body)
    body = re.sub(r"\bnrpy_c(\d+)\b", lambda m: "in_gfs[IDX4(%sGF, i0,i1,i2)]" % gf[m.group(1)], body)
    body = re.sub(r"\*NRPY_OWNED\(&(\w+\[IDX4\(\w+, i0, i1, i2\)\])\)", r"\1", body)
    return "\n".join(l for l in body.splitlines()
                     if not re.match(r"\s*const double (in_gfs\[|nrpy_h\d+ =)", l))


_TILE_ADDR = re.compile(r"\s*(#pragma (push|pop)_macro\(\"(IDX4|CCTK_GFINDEX3D)\"\)|#undef (IDX4|CCTK_GFINDEX3D)|"
                        r"#define (IDX4|CCTK_GFINDEX3D) \w+_TILE_X[CL]|NRPY_(ETK_)?TILE_ORIGIN\(\w+\);)$")


def _statements(kernel, bounds):
    """Statements after the bounds check, accesses reduced to ACC(GF, rhs?, offset). The
    lines that switch a kernel to tile-relative addressing (the same accesses, other
    address arithmetic) are dropped."""
    kernel = "\n".join(l for l in kernel.splitlines() if not _TILE_ADDR.match(l))
    body = kernel[kernel.index(bounds) + len(bounds):]
    if "nrpy_xshift<" in body:
        body = _undo_xshift(body)
    body = _SENR.sub(lambda m: "ACC(%s,%d,%s)" % (m.group(2), m.group(1) == "rhs_gfs",
                                                   m.group(3).replace(" ", "")), body)
    body = _ETK.sub(lambda m: "ACC(%s,%d,%s)" % (m.group(1).upper(), bool(m.group(2)),
                                                  m.group(3).replace(" ", "")), body)
    return [l.strip() for l in body.splitlines() if l.strip()]


PAIRS = [("Baikal_BSSN_Ricci_g0", "Ricci_eval_g0"), ("Baikal_BSSN_Ricci_g1", "Ricci_eval_g1"),
         ("Baikal_BSSN_RHSs_g0", "rhs_eval_g0"), ("Baikal_BSSN_RHSs_g1", "rhs_eval_g1"),
         ("Baikal_BSSN_RHSs_g2", "rhs_eval_g2"),
         ("Baikal_BSSN_constraints_H", "Hamiltonian_constraint"),
         ("Baikal_BSSN_constraints_MU", "momentum_constraint"),
         ("Baikal_enforce_detgammabar_constraint_k", "enforce_detgammabar_constraint")]


@pytest.mark.parametrize("etk_name,senr_name", PAIRS, ids=[p[0] for p in PAIRS])
def test_plugin_kernels_equal_senr_kernels(etk_name, senr_name):
    etk = open(os.path.join(GEN, "bssn_etk.hip")).read()
    senr = open(os.path.join(GEN, "bssn_cartesian.hip")).read()
    ke, ks = _kernel(etk, etk_name), _kernel(senr, senr_name)
    bounds_e = "if (i0 >= nrpy_hi0 || i1 >= nrpy_hi1 || i2 >= nrpy_hi2) return;"
    interior = "if (i0 >= Nxx0 + NGHOSTS || i1 >= Nxx1 + NGHOSTS || i2 >= Nxx2 + NGHOSTS) return;"
    allpts = "if (i0 >= Nxx_plus_2NGHOSTS0 || i1 >= Nxx_plus_2NGHOSTS1 || i2 >= Nxx_plus_2NGHOSTS2) return;"
    se = _statements(ke, bounds_e)
    ss = _statements(ks, interior if interior in ks else allpts)
    assert len(se) > 10
    if "nrpy_xshift<" in ks:
        # the lane-move form also moves the stencil-level statements next to their uses
        assert sorted(se) == sorted(ss)
    else:
        assert se == ss
    # the staged kernels stage the same gridfunctions in the same order
    assert ("nrpy_lds" in ke) == ("nrpy_lds" in ks)
    lds_e = [x.upper() for x in re.findall(r"copy the .*? of ([\w, ]+) to LDS", ke)]
    lds_s = [x.upper() for x in re.findall(r"copy the .*? of ([\w, ]+) to LDS", ks)]
    assert lds_e == lds_s


def test_constraint_kernels_cover_h_and_mu():
    etk = open(os.path.join(GEN, "bssn_etk.hip")).read()
    k = _kernel(etk, "Baikal_BSSN_constraints_H") + _kernel(etk, "Baikal_BSSN_constraints_MU")
    for name in ("HGF", "MU0GF", "MU1GF", "MU2GF"):
        assert re.search(r"\b%s\[CCTK_GFINDEX3D\(cctkGH, i0, i1, i2\)\] = " % name, k), name


@pytest.fixture(scope="module")
def baikal():
    from nrpytutorial_amd.etk import Baikal
    return Baikal()


def _call(b, gh, evol=None, rbar=None, FD_order=4):
    p = b.new_params()
    ev = b.pointers(evol if evol is not None else [0x1000] * len(b.evol), len(b.evol))
    rb = b.pointers(rbar if rbar is not None else [0x1000] * len(b.auxevol), len(b.auxevol))
    return b.lib.Baikal_driver_pt1_BSSN_Ricci(ctypes.byref(gh) if gh is not None else None, ctypes.byref(p),
                                              FD_order, ev, rb, None)


def test_host_side_errors(baikal):
    """Each case returns before any HIP call (so no GPU is needed)."""
    from nrpytutorial_amd.etk import cctk_grid
    b = baikal
    good = dict(lsh=(20, 18, 16), nghostzones=(3, 3, 3), delta_space=(0.1, 0.1, 0.1))
    assert _call(b, None) == -1
    assert _call(b, cctk_grid(**good), FD_order=6) == -4
    assert _call(b, cctk_grid(**good), FD_order=2) == -4
    assert _call(b, cctk_grid(**dict(good, nghostzones=(3, 2, 3)))) == -2     # stencils need 3
    assert _call(b, cctk_grid(**dict(good, lsh=(6, 18, 16)))) == -2          # no interior point
    assert _call(b, cctk_grid(**good, ash=(19, 18, 16))) == -2                # ash < lsh
    assert _call(b, cctk_grid(**dict(good, delta_space=(0.1, 0.0, 0.1)))) == -4
    assert _call(b, cctk_grid(**good, ash=(1 << 11, 1 << 10, 1 << 10))) == -2  # 2^31 points per array
    # xy-plane past the bound of the tile-addressed kernels (9.6e8 points, under 2^31)
    assert _call(b, cctk_grid(**good, ash=(20000, 3000, 16))) == -2
    evol = [0x1000] * len(b.evol)
    evol[5] = 0
    assert _call(b, cctk_grid(**good), evol=evol) == -1
    rbar = [0x1000] * len(b.auxevol)
    rbar[-1] = 0
    assert _call(b, cctk_grid(**good), rbar=rbar) == -1
    p = b.new_params()
    gh = cctk_grid(**good)
    assert b.lib.Baikal_driver_pt2_BSSN_RHSs(ctypes.byref(gh), ctypes.byref(p), 8, None, None, None, None) == -4
    assert b.lib.Baikal_BSSN_constraints(ctypes.byref(gh), ctypes.byref(p), 4, b.pointers([0] * 24, 24),
                                         b.pointers([0x1000] * 4, 4), None) == -1
    assert b.lib.Baikal_enforce_detgammabar_constraint(ctypes.byref(gh), None, b.pointers([0x1000] * 24, 24),
                                                       None) == -1


def test_pointer_order_is_the_bundle_order(baikal):
    from nrpytutorial_amd.runtime import load_meta
    m = load_meta("bssn_cartesian")
    assert baikal.evol == m["gfs"]["evol"]
    assert baikal.auxevol == m["gfs"]["auxevol"]
    assert baikal.aux == ["H", "MU0", "MU1", "MU2"]
    hdr = open(os.path.join(ROOT, "include", "nrpy_amd_bssn_etk.h")).read()
    listed = re.search(r"evol_gfs / rhs_gfs \(24\): (.*?) \(the thorn", hdr, re.S).group(1)
    listed = [w for w in listed.split() if w != "*"]
    assert listed == baikal.evol
