"""The tiled i0-face kernel of the Sommerfeld condition (nrpy_sommerfeld_xface_kernel in the
generated Cartesian bundles), on the CPU: everything its stencils read must be inside the
box a tile copies to LDS, or the kernel would read LDS it never wrote.

The bundle's own text is compiled with g++ (its paramstruct, the generated radial-derivative
function nrpy_sommerfeld_r_dfdr and the tile helpers), with an accessor that records each
read. For every tile and lane of a grid and every ghost layer the lane's row takes, at the
point and at its inward neighbour as nrpy_sommerfeld_rhs evaluates them: every read lies in
the tile's box, every LDS index is below NRPY_SOM_BOXN, and the value read through the box
equals the array's. The kernel takes each point of the i0-face groups whose (i1, i2) is
interior exactly once, and no other point; the rest of those groups (the edge rows, which
bcstruct_create lists per layer) is then every point whose (i1, i2) is not interior (the
groups as nrpy_set_bcstruct builds them when every ghost point is an outer point). Grids from
the smallest the condition accepts to sizes with partial tiles, in both bundles (radial FD
orders 2 and 4). Control: a box one point narrower along i1 misses reads.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "nrpytutorial_amd", "csrc", "generated")


def _pieces(bundle):
    src = open(os.path.join(GEN, bundle + ".hip")).read()
    out = []
    out.append(re.search(r"^#define NGHOSTS \d+\n", src, re.M).group(0))
    out.append(re.search(r"typedef struct __paramstruct__ \{.*?\} paramstruct;\n", src, re.S).group(0))
    out.append(re.search(r"^#define NRPY_SOMMERFELD_FD \d+\n", src, re.M).group(0))
    start = src.index("template <typename FA>\nstatic __device__ void nrpy_sommerfeld_r_dfdr(")
    out.append(src[start:src.index("\n}\n", start) + 3])
    start = src.index("#define NRPY_SOM_TY ")
    end = src.index("// Block (tile blockIdx.x of the interior rows", start)
    out.append(src[start:end])
    return "".join(out)


PROG = r"""
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <algorithm>
using std::min;
using std::max;
#define __device__
#define __host__
#define __restrict__
%(pieces)s
static long bad_box, bad_idx, bad_val, reads;
int main(int argc, char **argv) {
  const int Nx0 = atoi(argv[1]), Nx1 = atoi(argv[2]), Nx2 = atoi(argv[3]), narrow = atoi(argv[4]);
  paramstruct params;
  memset(&params, 0, sizeof(params));
  params.Nxx_plus_2NGHOSTS0 = Nx0 + 2 * NGHOSTS;
  params.Nxx_plus_2NGHOSTS1 = Nx1 + 2 * NGHOSTS;
  params.Nxx_plus_2NGHOSTS2 = Nx2 + 2 * NGHOSTS;
  params.invdx0 = params.invdx1 = params.invdx2 = 1.0;
  const int N0 = params.Nxx_plus_2NGHOSTS0, N1 = params.Nxx_plus_2NGHOSTS1, N2 = params.Nxx_plus_2NGHOSTS2;
  double *xx[3];
  for (int d = 0; d < 3; d++) {
    const int n = d == 0 ? N0 : d == 1 ? N1 : N2;
    xx[d] = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) xx[d][i] = -1.0 + 0.1 * i + 0.013 * d;
  }
  double *f = (double *)malloc(sizeof(double) * (size_t)N0 * N1 * N2);
  for (size_t q = 0; q < (size_t)N0 * N1 * N2; q++) f[q] = 1.0 + 1e-3 * (double)(q %% 9973);
  unsigned char *seen = (unsigned char *)calloc((size_t)N0 * N1 * N2, 1);
  static double box[NRPY_SOM_BOXN];
  long covered = 0;
  {
    const int nt = nrpy_som_xface_ntiles(N1, N2);
    for (int t = 0; t < nt; t++) {
      nrpy_som_xtile T = nrpy_som_xface_tile(N0, N1, N2, t);
      T.by1 -= narrow;
      const int BY = T.by1 - T.by0, BZ = T.bz1 - T.bz0, nbox = NRPY_SOM_XW * BY * BZ;
      if (nbox > NRPY_SOM_BOXN) { bad_idx++; continue; }
      for (int e = 0; e < nbox; e++) {
        const int x = e %% NRPY_SOM_XW, y = (e / NRPY_SOM_XW) %% BY, z = e / (NRPY_SOM_XW * BY);
        box[e] = f[(T.bx0 + x) + N0 * ((T.by0 + y) + N1 * (T.bz0 + z))];
      }
      for (int lane = 0; lane < NRPY_SOM_TY * NRPY_SOM_TZ; lane++) {
        const int i1 = T.y0 + lane %% NRPY_SOM_TY, i2 = T.z0 + lane / NRPY_SOM_TY;
        if (i1 >= T.hi1 || i2 >= T.hi2) continue;
        for (int gz = 0; gz < NGHOSTS; gz++) {
        const int i0 = T.F0 > 0 ? NGHOSTS - 1 - gz : N0 - NGHOSTS + gz;
        seen[i0 + (size_t)N0 * (i1 + (size_t)N1 * i2)]++;
        covered++;
        auto fa = [&](const int a, const int b, const int c) {
          reads++;
          if (a < T.bx0 || a >= T.bx0 + NRPY_SOM_XW || b < T.by0 || b >= T.by1 || c < T.bz0 || c >= T.bz1) {
            bad_box++;
            return 0.0;
          }
          const int li = (a - T.bx0) + NRPY_SOM_XW * ((b - T.by0) + BY * (c - T.bz0));
          if (li < 0 || li >= NRPY_SOM_BOXN) { bad_idx++; return 0.0; }
          if (box[li] != f[a + N0 * (b + N1 * c)]) bad_val++;
          return box[li];
        };
        double r, dfdr;
        for (int p = 0; p < 2; p++) {   /* the point, then its inward neighbour */
          const int j0 = i0 + p * T.F0;
          fa(j0, i1, i2);
          nrpy_sommerfeld_r_dfdr(params, xx[0], xx[1], xx[2], fa, T.F0, 0, 0, j0, i1, i2, &r, &dfdr);
        }
        }
      }
    }
  }
  /* the i0-face groups of a grid whose ghost points are all outer points: the kernel takes
     the points of interior rows once each, the edge rows (the rest) are left to the lists */
  long expect = 0, missed = 0, twice = 0;
  for (int gz = 0; gz < NGHOSTS; gz++)
    for (int side = 0; side < 2; side++)
      for (int i2 = NGHOSTS - gz; i2 < N2 - NGHOSTS + gz; i2++)
        for (int i1 = NGHOSTS - gz; i1 < N1 - NGHOSTS + gz; i1++) {
          const int i0 = side == 0 ? NGHOSTS - 1 - gz : N0 - NGHOSTS + gz;
          const size_t q = i0 + (size_t)N0 * (i1 + (size_t)N1 * i2);
          const unsigned char s = seen[q];
          seen[q] = 0;
          const int interior = i1 >= NGHOSTS && i1 < N1 - NGHOSTS && i2 >= NGHOSTS && i2 < N2 - NGHOSTS;
          if (!interior) { if (s != 0) twice++; continue; }
          expect++;
          if (s == 0) missed++;
          if (s > 1) twice++;
        }
  /* and no point outside the groups */
  for (size_t q = 0; q < (size_t)N0 * N1 * N2; q++)
    if (seen[q]) twice++;
  printf("%%ld %%ld %%ld %%ld %%ld %%ld %%ld %%ld\n", bad_box, bad_idx, bad_val, reads, covered, expect, missed, twice);
  return 0;
}
"""


@pytest.fixture(scope="module", params=["maxwell_cartesian", "scalarwave_cartesian"])
def exe(request, tmp_path_factory):
    tmp = tmp_path_factory.mktemp("xface")
    c = tmp / (request.param + ".cpp")
    c.write_text(PROG % {"pieces": _pieces(request.param)})
    x = tmp / request.param
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-w", "-include", "cstring", "-o", str(x), str(c)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return str(x)


def _run(exe, *a):
    r = subprocess.run([exe] + [str(v) for v in a], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return [int(v) for v in r.stdout.split()]


@pytest.mark.parametrize("grid", [(3, 3, 3), (5, 4, 3), (16, 12, 8), (17, 30, 9), (40, 33, 21)])
def test_xface_reads_stay_in_the_box_and_tiles_cover_the_faces(exe, grid):
    bad_box, bad_idx, bad_val, reads, covered, expect, missed, twice = _run(exe, *grid, 0)
    assert reads > 0 and covered == expect > 0
    assert (bad_box, bad_idx, bad_val, missed, twice) == (0, 0, 0, 0, 0)


def test_xface_control_narrow_box_misses_reads(exe):
    assert _run(exe, 17, 40, 9, 1)[0] > 0
