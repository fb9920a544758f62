"""The C ABI: every bundle library loads (no GPU needed to load it) and exports
every entry point that include/*.h declares for its system; the headers
compile as C and as C++; status codes and the block-shape setter behave."""
import ctypes
import glob
import json
import os
import re
import subprocess

import numpy as np
import pytest

from nrpytutorial_amd import runtime as rt
from nrpytutorial_amd.codegen.systems import BUNDLES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

FAMILY_HEADER = {"scalarwave": "nrpy_amd_scalarwave.h",
                 "scalarwave_curvilinear": "nrpy_amd_scalarwave_curvilinear.h",
                 "bssn": "nrpy_amd_bssn.h",
                 "maxwell": "nrpy_amd_maxwell.h",
                 "maxwell_curvilinear": "nrpy_amd_maxwell_curvilinear.h",
                 "bssn_etk": "nrpy_amd_bssn_etk.h"}
# Families whose library is only the plugin-form kernels (no grid set-up, boundary
# conditions or time stepping from nrpy_amd.h).
KERNELS_ONLY = {"bssn_etk": 7}


def declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^(?:int|void|double)\s+\**([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M)))


def family(bundle):
    meta = rt.load_meta(bundle)
    return meta.get("family", bundle.split("_")[0])


def test_every_header_is_used():
    used = {"nrpy_amd.h"} | set(FAMILY_HEADER.values())
    assert {os.path.basename(p) for p in glob.glob(os.path.join(INC, "*.h"))} == used


@pytest.mark.parametrize("bundle", BUNDLES)
def test_library_exports_declared_symbols(bundle):
    path = os.path.join(rt.LIBDIR, "libnrpy_%s.so" % bundle)
    lib = ctypes.CDLL(path)
    fam = family(bundle)
    if fam in KERNELS_ONLY:
        want = declared(FAMILY_HEADER[fam])
        assert len(want) == KERNELS_ONLY[fam]
    else:
        want = declared("nrpy_amd.h") + declared(FAMILY_HEADER[fam])
        assert len(want) > 20
    missing = [s for s in want if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_headers_compile(lang, compiler, tmp_path):
    src = tmp_path / ("t." + ("c" if lang == "c" else "cpp"))
    for h in FAMILY_HEADER.values():
        src.write_text('#include "%s"\nint main(void) { return NRPY_OK; }\n' % h)
        r = subprocess.run([compiler, "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I", INC, "-I", "/opt/rocm/include",
                            str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]


def test_status_codes_match_runtime():
    text = open(os.path.join(INC, "nrpy_amd.h")).read()
    codes = {int(v): k for k, v in re.findall(r"#define (NRPY_ERR_\w+) (-\d+)", text)}
    assert sorted(codes) == sorted(rt.NRPY_ERRORS)
    pre = open(os.path.join(ROOT, "nrpytutorial_amd", "csrc", "nrpy_runtime_pre.h")).read()
    for v, k in codes.items():
        assert re.search(r"#define %s %d\b" % (k, v), pre), k


def test_block_shape_setter_validates():
    b = rt.Bundle("bssn_cartesian")
    f = b.lib.nrpy_set_block_shape
    assert f(0, 64, 1, 4) == 0
    assert f(0, 0, 0, 0) == 0
    assert f(0, 64, 2, 4) == -4       # 512 threads
    assert f(0, 48, 1, 1) == -4       # not a multiple of 64
    assert f(256, 64, 1, 1) == -4     # beyond the table (NRPY_MAX_KERNELS = 256)
    assert f(-1, 64, 1, 1) == -4
    # every kernel id a bundle hands out fits the table, so tuning reaches every kernel
    for name in BUNDLES:
        ids = rt.load_meta(name).get("kernel_ids", {})
        assert all(0 <= k < 256 for k in ids.values()), name
    assert f(max(rt.load_meta("bssn_spherical_sym2")["kernel_ids"].values()), 64, 4, 1) == 0
    assert f(max(rt.load_meta("bssn_spherical_sym2")["kernel_ids"].values()), 0, 0, 0) == 0


def test_null_and_shape_errors_on_host():
    """Error paths that return before touching the device."""
    b = rt.Bundle("scalarwave_cartesian")
    L = b.lib
    n = (ctypes.c_int * 3)(8, 8, 8)
    xx = (ctypes.c_void_p * 3)()
    assert L.set_Nxx_dxx_invdx_params__and__xx(0, n, None, xx) == -1
    p = b.new_params()
    assert L.set_Nxx_dxx_invdx_params__and__xx(2, n, ctypes.byref(p), xx) == -4
    bad = (ctypes.c_int * 3)(8, 0, 8)
    assert L.set_Nxx_dxx_invdx_params__and__xx(0, bad, ctypes.byref(p), xx) == -2
    h = ctypes.c_void_p()
    assert L.bcstruct_create_host(None, xx, ctypes.byref(h)) == -1
    assert L.MoL_graph_launch(None, 1, None) == -1


def test_paramstruct_layout_matches_meta():
    """The ctypes mirror built from <bundle>.json has the size of the C struct
    (members sorted by name, NRPy_param_funcs.py:241-242)."""
    for bundle in BUNDLES:
        meta = rt.load_meta(bundle)
        names = [m[0] for m in meta["paramstruct"]]
        assert names == sorted(names)
        src = open(os.path.join(rt.GENDIR, bundle + ".hip")).read()
        body = src[src.index("typedef struct __paramstruct__ {"):]
        body = body[:body.index("} paramstruct;")]
        cnames = re.findall(r"^\s+(?:int|double) (\w+);", body, flags=re.M)
        assert cnames == names, bundle


def test_grid_size_limits_on_host():
    """A grid whose gridfunction exceeds 2^31 points, or whose Nxx_plus_2NGHOSTS does not
    match Nxx + 2 NGHOSTS, is rejected with NRPY_ERR_SHAPE before any launch (so no GPU
    is needed); the largest grid that fits passes the check."""
    b = rt.Bundle("scalarwave_cartesian")
    G = b.NGHOSTS
    fake = ctypes.c_void_p(0x1000)

    def call(n0, n1, n2, plus=None):
        p = b.new_params(Nxx0=n0, Nxx1=n1, Nxx2=n2)
        q = plus or (n0 + 2 * G, n1 + 2 * G, n2 + 2 * G)
        p.Nxx_plus_2NGHOSTS0, p.Nxx_plus_2NGHOSTS1, p.Nxx_plus_2NGHOSTS2 = q
        return b.lib.rhs_eval(ctypes.byref(p), fake, fake, None)

    assert call(1290, 1290, 1290) == -2               # 1296^3 > 2^31 points
    assert call(16, 16, 16, plus=(22, 22, 21)) == -2  # inconsistent ghost sizes
    assert call(0, 16, 16) == -2
    assert b.lib.rhs_eval(None, fake, fake, None) == -1


C_DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include "nrpy_amd_scalarwave_curvilinear.h"
%s
int main(void) {
  paramstruct params;
%s
  const int Nxx[3] = {16, 8, 12};
  double *xx[3];
  for (int d = 0; d < 3; d++) xx[d] = (double *)malloc(sizeof(double) * (Nxx[d] + 2 * 3));
  if (set_Nxx_dxx_invdx_params__and__xx(0, Nxx, &params, xx) != NRPY_OK) return 1;
  double x[3];
  xx_to_Cart(&params, xx, 5, 4, 7, x);
  printf("%%.17g %%.17g %%.17g %%.17g\n", find_timestep(&params, xx, 0.5), x[0], x[1], x[2]);
  if (set_Nxx_dxx_invdx_params__and__xx(2, Nxx, &params, xx) != NRPY_ERR_PARAM) return 2;
  return 0;
}
"""


def test_c_driver_links_and_matches_python(tmp_path):
    """The reference's drivers are C: a C main() built with gcc against the headers, the
    paramstruct written from <bundle>.json, and the bundle library runs the host-side
    grid setup (no GPU) and agrees with the ctypes path bit for bit."""
    bundle = "scalarwave_sinhsphericalv2"
    meta = rt.load_meta(bundle)
    struct = "struct __paramstruct__ {\n%s};\n" % "".join(
        "  %s %s;\n" % ("int" if t == "int" else "double", n) for n, t, _ in meta["paramstruct"])
    init = "".join("  params.%s = %s;\n" % (n, repr(int(d)) if t == "int" else repr(float(d)))
                   for n, t, d in meta["paramstruct"])
    src = tmp_path / "driver.c"
    src.write_text(C_DRIVER % (struct, init))
    exe = tmp_path / "driver"
    r = subprocess.run(["gcc", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", INC, "-I", "/opt/rocm/include", str(src),
                        "-o", str(exe), "-L", rt.LIBDIR, "-lnrpy_%s" % bundle, "-L/opt/rocm/lib", "-lamdhip64",
                        "-Wl,-rpath," + rt.LIBDIR + ":/opt/rocm/lib"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, (out.returncode, out.stderr[-2000:])
    got = [float(v) for v in out.stdout.split()]
    b = rt.Bundle(bundle)
    xx = [np.zeros(n + 2 * b.NGHOSTS) for n in (16, 8, 12)]
    p = b.new_params()
    ptrs = (ctypes.c_void_p * 3)(*[x.ctypes.data for x in xx])
    assert b.lib.set_Nxx_dxx_invdx_params__and__xx(0, (ctypes.c_int * 3)(16, 8, 12), ctypes.byref(p), ptrs) == 0
    x = (ctypes.c_double * 3)()
    b.lib.xx_to_Cart(ctypes.byref(p), ptrs, 5, 4, 7, x)
    want = [b.lib.find_timestep(ctypes.byref(p), ptrs, 0.5)] + list(x)
    assert got == want


@pytest.mark.parametrize("bundle", ["scalarwave_cartesian", "scalarwave_sinhsphericalv2"])
def test_find_timestep_divides_by_wavespeed(bundle):
    """find_timestep returns dsmin*CFL_FACTOR/wavespeed (reference_metric.py:1434): doubling the
    scalar wave's speed halves dt, in the bundle and in the oracle; BSSN has no wavespeed
    parameter and divides by the drivers' 1.0."""
    from oracle_lib import Oracle, OracleGrid
    b = rt.Bundle(bundle)
    Nxx = (16, 12, 10)
    dts = {}
    for c in (1.0, 2.0):
        p, xx = rt.host_grid(b, Nxx, False, wavespeed=c)
        dts[c] = b.lib.find_timestep(ctypes.byref(p), rt._xx_ptrs(xx), 0.5)
        og = OracleGrid(Oracle(bundle), Nxx, {"wavespeed": c})
        assert og.find_timestep(0.5) == pytest.approx(dts[c], rel=1e-15)
    assert dts[2.0] == pytest.approx(0.5 * dts[1.0], rel=1e-15)
    assert "wavespeed" not in [m[0] for m in rt.load_meta("bssn_cartesian")["paramstruct"]]
    src = open(os.path.join(rt.GENDIR, "bssn_cartesian.hip")).read()
    assert "return dsmin*CFL_FACTOR/1.0;" in src
