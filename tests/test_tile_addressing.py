"""Tile-relative addressing (codegen/addressing.py, NRPY_TILE_ORIGIN / IDX4_TILE_XC /
IDX4_TILE_XL and the toolkit-form NRPY_ETK_TILE_ORIGIN / CCTK_GFINDEX3D_TILE_XC / _XL in
nrpy_runtime_pre.h), on the CPU.

* The index arithmetic, by brute force: the macros' own text, compiled with gcc, gives
  the plain IDX4 index, and the toolkit form the generated CCTK_GFINDEX3D index, for
  every thread of every tile of a small grid (16x4x4 and 64x1x4 tiles, the grid's last
  tiles partial, ghost zones of NGHOSTS and deeper) and every offset within NGHOSTS along
  each axis, in both forms; also at sampled points of a grid whose indices exceed 2^32
  and of the widest xy-plane the bundle's parameter check accepts. Control: the same
  program with the tile's corner not moved back in y (a lane offset that can go
  negative) fails.
* The bound in the parameter check is exact enough: at the widest accepted plane the
  largest lane offset is below 2^32 bytes.
* The pass: it refuses accesses beyond NGHOSTS, unknown index forms and macros that
  index on their own; in the generated Cartesian BSSN bundle exactly the kernels named in
  the bundle's meta are rewritten, with the bound in the parameter check.
"""
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRE = os.path.join(ROOT, "nrpytutorial_amd", "csrc", "nrpy_runtime_pre.h")
GEN = os.path.join(ROOT, "nrpytutorial_amd", "csrc", "generated")


def _macro(src, name):
    m = re.search(r"^#define %s\b.*?(?<!\\)\n" % name, src, re.M | re.S)
    assert m, name
    return m.group(0)


def _macros():
    src = open(PRE).read()
    etk = open(os.path.join(GEN, "bssn_etk.hip")).read()
    return "".join(_macro(src, n) for n in (
        "IDX3S", "IDX4", "NRPY_TILE_ORIGIN_N", "NRPY_TILE_XC", "NRPY_TILE_XL", "NRPY_TILE_ORIGIN",
        "NRPY_GF_OFFSET", "IDX4_TILE_XC", "IDX4_TILE_XL", "NRPY_ETK_TILE_ORIGIN", "CCTK_GFINDEX3D_TILE_XC",
        "CCTK_GFINDEX3D_TILE_XL")) + _macro(etk, "CCTK_GFINDEX3D")


PROG = r"""
#include <stdio.h>
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
static struct { int x, y, z; } threadIdx;
typedef struct { int ash0, ash1; } geom;
#define NGHOSTS 3
%(macros)s
/* mismatches of both forms against IDX4 and CCTK_GFINDEX3D over the threads of every
   tile (sampled == 0) or of the tiles at both ends of each axis and the middle one along
   y; the interior starts lo >= NGHOSTS points in (the toolkit's ghost zones) */
static long check(int Nxx_plus_2NGHOSTS0, int Nxx_plus_2NGHOSTS1, int Nxx_plus_2NGHOSTS2,
                  int BX, int BY, int BZ, int sampled, int g, int lo) {
  const geom cctkGH = {Nxx_plus_2NGHOSTS0, Nxx_plus_2NGHOSTS1};
  const int n[3] = {Nxx_plus_2NGHOSTS0 - 2 * lo, Nxx_plus_2NGHOSTS1 - 2 * lo, Nxx_plus_2NGHOSTS2 - 2 * lo};
  const int nb[3] = {(n[0] + BX - 1) / BX, (n[1] + BY - 1) / BY, (n[2] + BZ - 1) / BZ};
  long bad = 0;
  for (int bz = 0; bz < nb[2]; bz++)
  for (int by = 0; by < nb[1]; by++)
  for (int bx = 0; bx < nb[0]; bx++) {
    if (sampled && !((bx == 0 || bx == nb[0] - 1) && (by == 0 || by == nb[1] - 1 || by == nb[1] / 2) &&
                     (bz == 0 || bz == nb[2] - 1)))
      continue;
    for (int tz = 0; tz < BZ; tz++)
    for (int ty = 0; ty < BY; ty++)
    for (int tx = 0; tx < BX; tx++) {
      threadIdx.x = tx; threadIdx.y = ty; threadIdx.z = tz;
      const int i0 = lo + bx * BX + tx, i1 = lo + by * BY + ty, i2 = lo + bz * BZ + tz;
      if (i0 >= n[0] + lo || i1 >= n[1] + lo || i2 >= n[2] + lo) continue;
      for (int dz = -NGHOSTS; dz <= NGHOSTS; dz++)
      for (int dy = -NGHOSTS; dy <= NGHOSTS; dy++)
      for (int dx = -NGHOSTS; dx <= NGHOSTS; dx++) {
        const size_t want = IDX4(g, i0 + dx, i1 + dy, i2 + dz);
        const size_t want_etk = CCTK_GFINDEX3D(cctkGH, i0 + dx, i1 + dy, i2 + dz);
        {
          NRPY_TILE_ORIGIN(0);
          if ((size_t)IDX4_TILE_XC(g, i0 + dx, i1 + dy, i2 + dz) != want) bad++;
        }
        {
          NRPY_TILE_ORIGIN(NGHOSTS);
          if ((size_t)IDX4_TILE_XL(g, i0 + dx, i1 + dy, i2 + dz) != want) bad++;
        }
        {
          NRPY_ETK_TILE_ORIGIN(0);
          if ((size_t)CCTK_GFINDEX3D_TILE_XC(cctkGH, i0 + dx, i1 + dy, i2 + dz) != want_etk) bad++;
        }
        {
          NRPY_ETK_TILE_ORIGIN(NGHOSTS);
          if ((size_t)CCTK_GFINDEX3D_TILE_XL(cctkGH, i0 + dx, i1 + dy, i2 + dz) != want_etk) bad++;
        }
      }
    }
  }
  return bad;
}
int main(int argc, char **argv) {
  int a[9];
  for (int q = 0; q < 9; q++) a[q] = atoi(argv[q + 1]);
  printf("%%ld\n", check(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]));
  return 0;
}
"""


@pytest.fixture(scope="module")
def progs(tmp_path_factory):
    d = tmp_path_factory.mktemp("tile_addr")
    out = {}
    good = _macros()
    # control: the tile's corner not moved back along y
    bad = good.replace("nrpy_o1 = i1 - (int)threadIdx.y - NGHOSTS", "nrpy_o1 = i1 - (int)threadIdx.y")
    assert bad.count("nrpy_o1 = i1 - (int)threadIdx.y,") == 1
    assert bad != good
    for name, macros in (("good", good), ("bad", bad)):
        c = d / (name + ".c")
        c.write_text(PROG % {"macros": macros})
        exe = d / name
        r = subprocess.run(["gcc", "-O2", "-o", str(exe), str(c)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        out[name] = str(exe)
    return out


def _run(exe, *args):
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return int(r.stdout)


@pytest.mark.parametrize("tile", [(16, 4, 4), (64, 1, 4)])
@pytest.mark.parametrize("grid", [(13, 9, 7), (70, 9, 8), (6, 5, 5)])
def test_tile_index_equals_idx4(progs, tile, grid):
    for lo in (3, 5):
        N = [n + 2 * lo for n in grid]
        for g in (0, 5):
            assert _run(progs["good"], *N, *tile, 0, g, lo) == 0


def test_tile_index_control_fails(progs):
    assert _run(progs["bad"], 19, 15, 13, 16, 4, 4, 0, 0, 3) > 0


def _plane_ok(N0, N1, tile):
    """False where the generated parameter check (addressing.bound_check_code) rejects
    the grid, its C expression evaluated for NGHOSTS = 3."""
    from nrpytutorial_amd.codegen import addressing
    saved = list(addressing.TILE_BOUND)
    try:
        addressing.TILE_BOUND[:] = list(tile)
        code = addressing.bound_check_code()
    finally:
        addressing.TILE_BOUND[:] = saved
    expr = (code.replace("(size_t)", "").replace("p->Nxx_plus_2NGHOSTS0", str(N0))
            .replace("p->Nxx_plus_2NGHOSTS1", str(N1)).replace("NGHOSTS", "3"))
    assert re.fullmatch(r"[\d\s()+*<>]+", expr), expr
    return not eval(expr)


def test_large_grids(progs):
    # indices beyond 2^32 (gridfunction 60 of a 4006 x 4006 x 36 grid)
    assert _run(progs["good"], 4006, 4006, 36, 16, 4, 4, 1, 60, 3) == 0
    assert _run(progs["good"], 4006, 4006, 36, 64, 1, 4, 1, 60, 3) == 0
    # the widest xy-plane the parameter check accepts for the bundle's tiles (64, 4, 4)
    tile = (64, 4, 4)
    N0 = 7000
    N1 = max(n for n in range(1000, 20000) if _plane_ok(N0, n, tile))
    assert not _plane_ok(N0, N1 + 1, tile)
    # largest lane offset in bytes: corner-relative point (BX - 1 + 2 NG, BY - 1 + 2 NG, BZ - 1 + 2 NG)
    top = 8 * ((tile[0] - 1 + 6) + N0 * ((tile[1] - 1 + 6) + N1 * (tile[2] - 1 + 6)))
    assert top < 1 << 32
    for t in ((16, 4, 4), (64, 1, 4)):
        assert _run(progs["good"], N0, N1, 12, *t, 1, 3, 3) == 0


def test_bound_check_code_matches():
    from nrpytutorial_amd.codegen import addressing
    saved = list(addressing.TILE_BOUND)
    try:
        addressing.TILE_BOUND[:] = [64, 4, 4]
        code = addressing.bound_check_code()
    finally:
        addressing.TILE_BOUND[:] = saved
    assert code == ("(size_t)(64 + 2*NGHOSTS) + (size_t)p->Nxx_plus_2NGHOSTS0 * ((4 + 2*NGHOSTS) + "
                    "(size_t)p->Nxx_plus_2NGHOSTS1 * (4 + 2*NGHOSTS)) > ((size_t)1 << 29)")


KERNEL = """__global__ void __launch_bounds__(256) k__gpu(const paramstruct params, const double *__restrict__ in_gfs, double *__restrict__ rhs_gfs) {
  const int i0 = NGHOSTS + nrpy_bx*(int)blockDim.x + (int)threadIdx.x;
  const int i1 = NGHOSTS + nrpy_by*(int)blockDim.y + (int)threadIdx.y;
  const int i2 = NGHOSTS + nrpy_bz*(int)blockDim.z + (int)threadIdx.z;
  if (i0 >= Nxx0 + NGHOSTS || i1 >= Nxx1 + NGHOSTS || i2 >= Nxx2 + NGHOSTS) return;
  {
    const double a = in_gfs[IDX4(UUGF, %s)];
    rhs_gfs[IDX4(UUGF, i0, i1, i2)] = a;
  }
}
"""


@pytest.mark.parametrize("acc,ok", [("i0-3,i1,i2+3", True), ("i0,i1-3,i2", True),
                                    ("nrpy_xl < 3 ? i0 - 3 : i0 + 3,i1,i2", True),
                                    ("i0,i1-4,i2", False), ("i0+4,i1,i2", False),
                                    ("i0,i1,i2+j", False), ("i1,i0,i2", False),
                                    ("nrpy_xl < 3 ? i0 - 2 : i0 + 3,i1,i2", False)])
def test_pass_refuses_accesses_out_of_range(acc, ok):
    from nrpytutorial_amd.codegen import addressing
    src = KERNEL % acc
    saved = list(addressing.TILE_BOUND), dict(addressing.APPLIED)
    try:
        if ok:
            out = addressing.tile_address(src, "xl", (16, 4, 4), 3)
            lines = out.splitlines()
            k = next(i for i, l in enumerate(lines) if "const int i2 = " in l)
            assert lines[k + 1:k + 5] == ["  NRPY_TILE_ORIGIN(NGHOSTS);", '#pragma push_macro("IDX4")',
                                          "#undef IDX4", "#define IDX4 IDX4_TILE_XL"]
            assert lines[-2:] == ['#pragma pop_macro("IDX4")', "}"]
        else:
            with pytest.raises(ValueError):
                addressing.tile_address(src, "xc", (16, 4, 4), 3)
        # VJP_G_AT with literal offsets within NGHOSTS is accepted, beyond it refused
        vjp = "const double b = VJP_G_AT(7, %s);\n    const double a ="
        if ok:
            addressing.tile_address(src.replace("const double a =", vjp % "0, -3, 2"), "xc", (16, 4, 4), 3)
        with pytest.raises(ValueError):
            addressing.tile_address(src.replace("const double a =", vjp % "4, 0, 0"), "xc", (16, 4, 4), 3)
        with pytest.raises(ValueError):
            addressing.tile_address(src.replace("const double a =", vjp % "0, k, 0"), "xc", (16, 4, 4), 3)
    finally:
        addressing.TILE_BOUND[:] = saved[0]
        addressing.APPLIED.clear()
        addressing.APPLIED.update(saved[1])


@pytest.mark.parametrize("bundle", ["bssn_cartesian", "bssn_etk"])
def test_bundle_kernels_with_tile_addressing(bundle):
    src = open(os.path.join(GEN, bundle + ".hip")).read()
    meta = json.load(open(os.path.join(GEN, bundle + ".json")))
    named = meta["tile_addressing"]
    assert named and set(named.values()) <= {"xc", "xl"}
    access, origin = ("IDX4", "NRPY_TILE_ORIGIN") if bundle == "bssn_cartesian" else \
        ("CCTK_GFINDEX3D", "NRPY_ETK_TILE_ORIGIN")
    kernels = {}
    for m in re.finditer(r"__global__ void [^\n]*? (\w+)__gpu\(", src):
        end = src.index("\n}\n", m.end())
        kernels[m.group(1)] = src[m.end():end]
    rewritten = {n: ("xc" if "%s(0)" % origin in b else "xl")
                 for n, b in kernels.items() if "%s(" % origin in b}
    assert rewritten == named
    for n, form in named.items():
        b = kernels[n]
        assert b.count('#pragma push_macro("%s")' % access) == b.count('#pragma pop_macro("%s")' % access) == 1
        assert "#define %s %s_TILE_%s" % (access, access, form.upper()) in b
    if bundle == "bssn_cartesian":
        # every launch of a rewritten kernel uses a fixed block shape (not the tunable table)
        tiles = []
        for n in named:
            launcher = re.search(r'extern "C" int %s\((.*?)\n}\n' % n, src, re.S)
            assert launcher, n
            block = re.search(r"const dim3 block = dim3\((\d+), (\d+), (\d+)\);", launcher.group(1))
            assert block and "nrpy_block(" not in launcher.group(1), n
            tiles.append([int(x) for x in block.groups()])
        bx, by, bz = (max(t[d] for t in tiles) for d in range(3))
        assert ("if ((size_t)(%d + 2*NGHOSTS) + (size_t)p->Nxx_plus_2NGHOSTS0 * ((%d + 2*NGHOSTS) + "
                "(size_t)p->Nxx_plus_2NGHOSTS1 * (%d + 2*NGHOSTS)) > ((size_t)1 << 29)) return 1;"
                % (bx, by, bz)) in src
    else:
        assert re.search(r"#define NRPY_ETK_TILE_TOO_WIDE\(A0, A1\) \(\(size_t\)\(\d+ \+ 2\*NGHOSTS\) \+ "
                         r"\(size_t\)\(A0\) \* .*\(size_t\)\(A1\) \* .*> \(\(size_t\)1 << 29\)\)", src)
        assert "if (interior && NRPY_ETK_TILE_TOO_WIDE((size_t)out->ash0, (size_t)out->ash1)) " \
               "return NRPY_ERR_SHAPE;" in src


def test_entry_points_refuse_too_wide_a_plane():
    """A consistent grid whose xy-plane is past the tile-addressing bound (7000 x 7700
    allocated, 3 planes deep: 1.6e8 points, under the 2^31 limit) is refused with
    NRPY_ERR_SHAPE by the host-side check of every Cartesian BSSN entry point that
    launches a rewritten kernel, before any HIP call."""
    import ctypes
    from nrpytutorial_amd.runtime import Bundle
    b = Bundle("bssn_cartesian")
    G = b.NGHOSTS
    N0, N1 = 7000, 7700
    assert not _plane_ok(N0, N1, (64, 4, 4)) and N0 * N1 * 3 < 2 ** 31
    p = b.new_params(Nxx0=N0 - 2 * G, Nxx1=N1 - 2 * G, Nxx2=1, Nxx_plus_2NGHOSTS0=N0, Nxx_plus_2NGHOSTS1=N1,
                     Nxx_plus_2NGHOSTS2=1 + 2 * G)
    dummy = ctypes.c_void_p(0x1000)     # never dereferenced: the check returns first
    lib = b.lib
    lib.rhs_eval.restype = ctypes.c_int
    assert lib.rhs_eval(None, ctypes.byref(p), dummy, dummy, dummy, None) == -2
    assert lib.Ricci_eval(None, ctypes.byref(p), dummy, dummy, None) == -2
    for k in range(3):
        f = getattr(lib, "rhs_eval_rk4_substep_g%d" % k)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 7 + [ctypes.c_double, ctypes.c_void_p]
        assert f(1, None, ctypes.byref(p), dummy, dummy, dummy, dummy, dummy, 0.1, None) == -2, k
