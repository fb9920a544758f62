"""The pass-2 helper VJP_G_AT of the generated vector-Jacobian products
(nrpy_runtime_pre.h), on the CPU. Its load is unconditional, so its address must lie
inside slot k's own array for every point of the grid and every offset within NGHOSTS,
and the selected value must be slot k at the offset point inside the interior, 0
outside. Checked by brute force with the macro's own text compiled by gcc (the array
read replaced by a recording load), for the plain index and for both tile-relative
forms, on grids down to one interior point per axis, with NGHOSTS 2 and 3. Control:
without the plane clamp, addresses leave the array.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRE = os.path.join(ROOT, "nrpytutorial_amd", "csrc", "nrpy_runtime_pre.h")
CLAMP = "min(max((i2) + (o2), 1), Nxx_plus_2NGHOSTS2 - 2)"


def _macro(src, name):
    m = re.search(r"^#define %s\b.*?(?<!\\)\n" % name, src, re.M | re.S)
    assert m, name
    return m.group(0)


def _macros(clamp=True):
    src = open(PRE).read()
    vjp = _macro(src, "VJP_G_AT")
    assert "vjp_G[IDX4(" in vjp and CLAMP in vjp
    # the array read becomes a call of the recording load; the plain IDX4 keeps its own name
    vjp = vjp.replace("vjp_G[IDX4(", "vjp_load(IDX(").replace(CLAMP + ")])", CLAMP + ")))")
    if not clamp:
        vjp = vjp.replace(CLAMP, "(i2) + (o2)")
    idx4 = _macro(src, "IDX4").replace("#define IDX4(", "#define IDX4_PLAIN(")
    return (_macro(src, "IDX3S") + idx4 + vjp + "".join(_macro(src, n) for n in (
        "NRPY_TILE_ORIGIN_N", "NRPY_TILE_XC", "NRPY_TILE_XL", "NRPY_TILE_ORIGIN", "NRPY_GF_OFFSET",
        "IDX4_TILE_XC", "IDX4_TILE_XL")))


PROG = r"""
#include <stdio.h>
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
static inline int min(int a, int b) { return a < b ? a : b; }
static inline int max(int a, int b) { return a > b ? a : b; }
static struct { int x, y, z; } threadIdx;
#define NGHOSTS %(G)d
static double nrpy_vjp_select(const int inside, const double v) { return inside ? v : 0.0; }
static size_t lo, hi;          /* slot k's array: flat indices [lo, hi) */
static long outside;
static double vjp_load(const size_t idx) {
  if (idx < lo || idx >= hi) outside++;
  return (double)idx;
}
%(macros)s
#define BX 16
#define BY 4
#define BZ 4
static long run(const int Nxx0, const int Nxx1, const int Nxx2, const int k, const int form) {
  const int Nxx_plus_2NGHOSTS0 = Nxx0 + 2 * NGHOSTS, Nxx_plus_2NGHOSTS1 = Nxx1 + 2 * NGHOSTS,
            Nxx_plus_2NGHOSTS2 = Nxx2 + 2 * NGHOSTS;
  const size_t Ntot = (size_t)Nxx_plus_2NGHOSTS0 * Nxx_plus_2NGHOSTS1 * Nxx_plus_2NGHOSTS2;
  lo = (size_t)k * Ntot;
  hi = lo + Ntot;
  outside = 0;
  long bad = 0;
  for (int i2 = 0; i2 < Nxx_plus_2NGHOSTS2; i2++)
  for (int i1 = 0; i1 < Nxx_plus_2NGHOSTS1; i1++)
  for (int i0 = 0; i0 < Nxx_plus_2NGHOSTS0; i0++) {
    threadIdx.x = i0 %% BX; threadIdx.y = i1 %% BY; threadIdx.z = i2 %% BZ;
    for (int o2 = -NGHOSTS; o2 <= NGHOSTS; o2++)
    for (int o1 = -NGHOSTS; o1 <= NGHOSTS; o1++)
    for (int o0 = -NGHOSTS; o0 <= NGHOSTS; o0++) {
      const int j0 = i0 + o0, j1 = i1 + o1, j2 = i2 + o2;
      const int in = j0 >= NGHOSTS && j0 < Nxx0 + NGHOSTS && j1 >= NGHOSTS && j1 < Nxx1 + NGHOSTS &&
                     j2 >= NGHOSTS && j2 < Nxx2 + NGHOSTS;
      const double want =
          in ? (double)(lo + j0 + (size_t)Nxx_plus_2NGHOSTS0 * (j1 + (size_t)Nxx_plus_2NGHOSTS1 * j2)) : 0.0;
      double got;
      if (form == 0) {
#define IDX IDX4_PLAIN
        got = VJP_G_AT(k, o0, o1, o2);
#undef IDX
      } else if (form == 1) {
        NRPY_TILE_ORIGIN(0);
#define IDX IDX4_TILE_XC
        got = VJP_G_AT(k, o0, o1, o2);
#undef IDX
      } else {
        NRPY_TILE_ORIGIN(NGHOSTS);
#define IDX IDX4_TILE_XL
        got = VJP_G_AT(k, o0, o1, o2);
#undef IDX
      }
      if (got != want) bad++;
    }
  }
  return bad + 1000000 * outside;
}
int main(int argc, char **argv) {
  (void)argc;
  printf("%%ld\n", run(atoi(argv[1]), atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5])));
  return 0;
}
"""


def _build(tmp, G, macros, name):
    c = tmp / (name + ".c")
    c.write_text(PROG % {"G": G, "macros": macros})
    exe = tmp / name
    r = subprocess.run(["gcc", "-O1", "-o", str(exe), str(c)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


@pytest.fixture(scope="module")
def exes(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("vjp_select")
    out = {G: _build(tmp, G, _macros(), "good%d" % G) for G in (2, 3)}
    out["bad"] = _build(tmp, 3, _macros(clamp=False), "bad")
    return out


def _run(exe, *a):
    r = subprocess.run([exe] + [str(x) for x in a], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return int(r.stdout)


@pytest.mark.parametrize("G", [2, 3])
@pytest.mark.parametrize("grid", [(1, 1, 1), (5, 3, 2), (13, 9, 7)])
@pytest.mark.parametrize("form", [0, 1, 2])
def test_vjp_g_at_stays_in_its_slot_and_selects(exes, G, grid, form):
    for k in (0, 4):
        assert _run(exes[G], *grid, k, form) == 0


def test_vjp_g_at_control_without_clamp_leaves_the_array(exes):
    assert _run(exes["bad"], 5, 3, 2, 0, 0) >= 1000000
