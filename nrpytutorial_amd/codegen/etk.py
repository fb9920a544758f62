"""Einstein Toolkit plugin form of the BSSN kernels (bundle ``bssn_etk``).

The reference generates the same BSSN kernels a second time for its Einstein
Toolkit thorn, Baikal: Cartesian coordinates, ``grid::GridFuncMemAccess = "ETK"``
(grid.py:92-100), one scheduled function per kernel
(BaikalETK/BaikalETK_C_kernels_codegen.py:264-330 for the Ricci tensor and the
RHSs; BaikalETK/BaikalETK_C_drivers_codegen.py:520-700 for the drivers that
pick the finite-difference order, the det(gammabar) enforcement and the
constraints). In that form every gridfunction is its own array, named as the
toolkit declares it (``hDD00GF``, right-hand sides ``hDD00_rhsGF``), indexed
with ``CCTK_GFINDEX3D(cctkGH, i, j, k)`` over the toolkit's allocated shape
``cctk_ash``; the loops run over ``[cctk_nghostzones, cctk_lsh -
cctk_nghostzones)``, and 1/dx comes from ``CCTK_DELTA_SPACE``.

Here the kernels are the same HIP kernels as the SENR-layout bundle: the same
symbolic expressions, output groups and finite-difference text, with only the
memory access spelled by ``gfaccess`` in ETK mode, and the mixed-derivative
stencils staged through LDS tiles the same way (codegen/lds.py handles both
access forms). The host entry points take what a scheduled function sees in
``CCTK_ARGUMENTS``: the grid shape (``nrpy_cctkGH``) and the gridfunction
pointers, as arrays in the gridfunction order of the bundle's metadata.
Finite-difference order 4 is built (the Start-to-Finish order); the drivers
reject other orders with NRPY_ERR_PARAM.
"""
import re
import sys
import time

from .. import NRPy_param_funcs as par
from .. import grid as gri
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from .. import finite_difference as fin
from .. import loop as lp
from ..outputC import lhrh
from . import bundle as bd
from . import lds
from . import addressing

FD_ORDER = 4

GEOMETRY = r"""
/* What a scheduled function of the Einstein Toolkit reads from CCTK_ARGUMENTS for
 * these kernels: local shape with ghost zones (cctk_lsh), allocated shape, i.e. the
 * memory strides (cctk_ash >= cctk_lsh), ghost zones per direction and
 * CCTK_DELTA_SPACE. */
typedef struct {
  int cctk_lsh[3];
  int cctk_ash[3];
  int cctk_nghostzones[3];
  double cctk_delta_space[3];
} nrpy_cctkGH;

/* The same, as the kernels take it: loop range [lo, hi), array extent lsh, strides. */
typedef struct {
  int lo[3], hi[3], lsh[3];
  int ash0, ash1;
  double invdx0, invdx1, invdx2;
} nrpy_etk_geom;

#define CCTK_GFINDEX3D(GH, i, j, k) ((i) + (GH).ash0 * ((j) + (GH).ash1 * (k)))

/* Every gridfunction pointer of a scheduled call: EVOL, their right-hand sides, AUXEVOL,
 * AUX. Each kernel takes the arrays it touches as its own parameters. */
typedef struct {
  double *p[NRPY_ETK_NPTRS];
} nrpy_etk_gfs;

/* Checks the grid of a scheduled call and converts it; interior = 0 loops over every
 * point (the det(gammabar) enforcement). */
static int nrpy_etk_geometry(const nrpy_cctkGH *g, const int interior, nrpy_etk_geom *out) {
  if (!g) return NRPY_ERR_NULL;
  size_t total = 1;
  for (int d = 0; d < 3; d++) {
    const int lsh = g->cctk_lsh[d], ng = g->cctk_nghostzones[d];
    if (lsh < 1 || g->cctk_ash[d] < lsh || ng < 0) return NRPY_ERR_SHAPE;
    /* the stencils reach NGHOSTS points from every interior point */
    if (interior && (ng < NGHOSTS || lsh - 2 * ng < 1)) return NRPY_ERR_SHAPE;
    if (!(g->cctk_delta_space[d] > 0.0)) return NRPY_ERR_PARAM;
    total *= (size_t)g->cctk_ash[d];
    out->lo[d] = interior ? ng : 0;
    out->hi[d] = interior ? lsh - ng : lsh;
    out->lsh[d] = lsh;
  }
  /* 32-bit point index inside one gridfunction */
  if (total >= ((size_t)1 << 31)) return NRPY_ERR_SHAPE;
  out->ash0 = g->cctk_ash[0];
  out->ash1 = g->cctk_ash[1];
  /* kernels with tile-relative addressing: 32-bit lane offsets within a plane */
  if (interior && NRPY_ETK_TILE_TOO_WIDE((size_t)out->ash0, (size_t)out->ash1)) return NRPY_ERR_SHAPE;
  out->invdx0 = 1.0 / g->cctk_delta_space[0];
  out->invdx1 = 1.0 / g->cctk_delta_space[1];
  out->invdx2 = 1.0 / g->cctk_delta_space[2];
  return NRPY_OK;
}
"""

# Tile-relative addressing (codegen/addressing.py) of the RHS group kernels, by the
# group's outputs. Measured at 192^3, one RHS evaluation (the three groups): 5.27 ms
# without; 5.03 with the aDD,trK group in form "xc"; 4.55 with the hDD,alpha,cf group in
# "xc" as well (it builds no x-neighbours by lane moves here). The Lambda-bar group is
# left as it is: 4.63 ms with it in "xc", 5.38 in "xl".
ETK_TILE_ADDRESS_OUTPUTS = (
    ({"aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22", "trK"}, "xc"),
    ({"hDD00", "hDD01", "hDD02", "hDD11", "hDD12", "hDD22", "alpha", "cf"}, "xc"),
)

_ACCESS = re.compile(r"\b([A-Za-z]\w*?)(_rhs)?GF\[CCTK_GFINDEX3D\(cctkGH,")
_WRITE = re.compile(r"\b(\w+GF)\[CCTK_GFINDEX3D\(cctkGH,[^\]]*\)\]\s*=(?!=)")
_GRID_PARAMS = re.compile(r"^(invdx|dxx|Nxx|xx)\d|^Nxx_plus_2NGHOSTS\d$")


class PointerTable:
    """Slot of every gridfunction pointer in nrpy_etk_gfs: EVOL, EVOL right-hand sides,
    AUXEVOL, AUX, each in the bundle's (alphabetical) gridfunction order."""

    def __init__(self):
        self.evol, self.aux, self.auxevol = gri.gridfunction_lists()
        self.n_evol, self.n_auxevol, self.n_aux = len(self.evol), len(self.auxevol), len(self.aux)
        self.size = 2 * self.n_evol + self.n_auxevol + self.n_aux

    def slot(self, name, rhs=False):
        if rhs:
            return self.n_evol + self.evol.index(name)
        t = gri.find_gftype(name)
        if t == "EVOL":
            return self.evol.index(name)
        if t == "AUXEVOL":
            return 2 * self.n_evol + self.auxevol.index(name)
        return 2 * self.n_evol + self.n_auxevol + self.aux.index(name)


def _pointer_params(body, table):
    """Kernel parameters ``hDD00GF`` / ``hDD00_rhsGF`` for every array the body touches
    (const unless the body assigns to it; __restrict__ on parameters is what tells the
    compiler that the arrays do not overlap) and their slots in nrpy_etk_gfs."""
    written = set(_WRITE.findall(body))
    out, seen = [], set()
    for name, rhs in _ACCESS.findall(body):
        ptr = name + ("_rhsGF" if rhs else "GF")
        if ptr in seen:
            continue
        seen.add(ptr)
        slot = table.slot(name, bool(rhs))
        out.append(("%sdouble *__restrict__ %s" % ("" if ptr in written else "const ", ptr), slot))
    return out


def etk_kernel(name, desc, body, table, interior=True):
    """``__global__`` kernel in the toolkit's access form; returns (kernel text, (block
    shape, pointer slots of its array parameters))."""
    used = [(n, t) for n, t in lp._used_cparams(body) if not _GRID_PARAMS.match(n)]
    head = ["const int nrpy_lo0 = cctkGH.lo[0], nrpy_lo1 = cctkGH.lo[1], nrpy_lo2 = cctkGH.lo[2];",
            "const int nrpy_hi0 = cctkGH.hi[0], nrpy_hi1 = cctkGH.hi[1], nrpy_hi2 = cctkGH.hi[2];",
            "const int nrpy_lsh0 = cctkGH.lsh[0], nrpy_lsh1 = cctkGH.lsh[1], nrpy_lsh2 = cctkGH.lsh[2];"]
    for d in range(3):
        if re.search(r"\binvdx%d\b" % d, body):
            head.append("const double invdx%d = cctkGH.invdx%d;" % (d, d))
    head += ["const %s %s = params.%s;" % (lp._ctype(t), n, n) for n, t in used]
    ptrs = _pointer_params(body, table)
    head.append("NRPY_BLOCK_COORDS(nrpy_hi0 - nrpy_lo0, nrpy_hi1 - nrpy_lo1, nrpy_hi2 - nrpy_lo2);")
    for d, t in ((0, "x"), (1, "y"), (2, "z")):
        head.append("const int i%d = nrpy_lo%d + nrpy_b%s*(int)blockDim.%s + (int)threadIdx.%s;" % (d, d, t, t, t))
    head.append("if (i0 >= nrpy_hi0 || i1 >= nrpy_hi1 || i2 >= nrpy_hi2) return;")
    text = ("/* %s */\n__global__ void __launch_bounds__(%s) %s__gpu(const nrpy_etk_geom cctkGH, "
            "const paramstruct params, %s) {\n%s}\n"
            % (desc, lp.launch_bounds(name), name, ", ".join(d for d, _ in ptrs),
               "".join("  " + l + "\n" for l in head + body.splitlines())))
    block = "dim3(%d, %d, %d)" % (lp.BLOCK_X, lp.BLOCK_Y, lp.BLOCK_Z)
    if interior and lp.LDS_STAGING:
        tile = lp.lds_tile(name)
        staged = lds.stage_kernel(text, lp.LDS_STAGING, tile)
        if staged is not None:
            text, block = staged, "dim3(%d, %d, %d)" % tile
            lp.LDS_KERNELS[name] = list(tile)
        else:
            tile = (lp.BLOCK_X, lp.BLOCK_Y, lp.BLOCK_Z)     # the launch's fixed block
        form = addressing.form_for(name)
        if form:
            # LDS_STAGING is the bundle's NGHOSTS (FD order / 2 + 1)
            text = addressing.tile_address(text, form, tile, lp.LDS_STAGING, name, access="CCTK_GFINDEX3D")
    return text, (block, [slot for _, slot in ptrs])


def _launch(name, launch):
    block, slots = launch
    return ("  {\n    const dim3 block = %s;\n"
            "    const dim3 grid(nrpy_nblocks(block, geom.hi[0] - geom.lo[0], geom.hi[1] - geom.lo[1], "
            "geom.hi[2] - geom.lo[2]));\n"
            "    hipLaunchKernelGGL(%s__gpu, grid, block, 0, stream, geom, *params, %s);\n"
            "    if (hipGetLastError() != hipSuccess) return NRPY_ERR_LAUNCH;\n  }\n"
            % (block, name, ", ".join("gfs.p[%d]" % q for q in slots)))


def _fill(arg, offset, count):
    """Copy a caller's pointer array into the table; every pointer must be non-NULL."""
    return ("  for (int g = 0; g < %d; g++) {\n"
            "    if (!%s[g]) return NRPY_ERR_NULL;\n"
            "    gfs.p[%d + g] = (double *)%s[g];\n  }\n" % (count, arg, offset, arg))


def entry_point(name, desc, args, fills, kernels, interior=True):
    """``extern "C"`` scheduled function: grid check, pointer table, one launch per kernel."""
    sig = "const nrpy_cctkGH *cctkGH, const paramstruct *params, " + ", ".join(args) + ", hipStream_t stream"
    body = ("  if (!params) return NRPY_ERR_NULL;\n"
            "  nrpy_etk_geom geom;\n  int rc = nrpy_etk_geometry(cctkGH, %d, &geom);\n  if (rc) return rc;\n"
            "  nrpy_etk_gfs gfs;\n  memset(&gfs, 0, sizeof(gfs));\n" % (1 if interior else 0))
    body += "".join(_fill(*f) for f in fills)
    body += "".join(_launch(k, b) for k, b in kernels)
    return ('/* %s */\nextern "C" int %s(%s) {\n%s  return NRPY_OK;\n}\n' % (desc, name, sig, body),
            'extern "C" int %s(%s);' % (name, sig))


def bssn_etk(bundle_name="bssn_etk"):
    from ..BSSN import BSSN_quantities as Bq, BSSN_RHSs as rhs, BSSN_gauge_RHSs as gaugerhs
    from ..BSSN import Enforce_Detgammabar_Constraint as EGC, BSSN_constraints as bssncon
    from .systems import rhs_groups, ricci_groups, lds_tile_for_group, MIN_TWO_WAVES_OUTPUTS
    t0 = time.time()
    par.set_parval_from_str("grid::DIM", 3)
    par.set_parval_from_str("finite_difference::FD_CENTDERIVS_ORDER", FD_ORDER)
    par.set_parval_from_str("reference_metric::CoordSystem", "Cartesian")
    par.set_parval_from_str("indexedexp::symmetry_axes", "")
    from .systems import BSSN_DIFFERENCED_GFS
    par.set_parval_from_str("finite_difference::FD_difference_from_centre", BSSN_DIFFERENCED_GFS)
    rfm.reference_metric()
    par.set_parval_from_str("BSSN.BSSN_gauge_RHSs::ShiftEvolutionOption", "GammaDriving2ndOrder_Covariant")
    par.set_parval_from_str("reference_metric::enable_rfm_precompute", "True")
    rfm.reference_metric()
    par.set_parval_from_str("BSSN.BSSN_quantities::LeaveRicciSymbolic", "True")
    rhs.BSSN_RHSs()
    gaugerhs.BSSN_gauge_RHSs()
    Bq.BSSN_basic_tensors()
    betaU = Bq.betaU
    par.set_parval_from_str("BSSN.BSSN_quantities::LeaveRicciSymbolic", "False")
    Bq.RicciBar__gammabarDD_dHatD__DGammaUDD__DGammaU()
    RbarDD = Bq.RbarDD
    gri.register_gridfunctions("AUX", "H")
    ixp.register_gridfunctions_for_single_rank1("AUX", "MU")
    bssncon.BSSN_constraints(add_T4UUmunu_source_terms=False)
    print("(BENCH) symbolic expressions %.1fs" % (time.time() - t0), flush=True)

    # From here on every gridfunction access is spelled the toolkit's way.
    par.set_parval_from_str("grid::GridFuncMemAccess", "ETK")
    detg_exprs = EGC.Enforce_Detgammabar_Constraint_symb_expressions()
    table = PointerTable()
    lhs_names = ["alpha", "cf", "trK"]
    rhs_exprs = [gaugerhs.alpha_rhs, rhs.cf_rhs, rhs.trK_rhs]
    for i in range(3):
        lhs_names += ["betU%d" % i, "lambdaU%d" % i, "vetU%d" % i]
        rhs_exprs += [gaugerhs.bet_rhsU[i], rhs.lambda_rhsU[i], gaugerhs.vet_rhsU[i]]
        for j in range(i, 3):
            lhs_names += ["aDD%d%d" % (i, j), "hDD%d%d" % (i, j)]
            rhs_exprs += [rhs.a_rhsDD[i][j], rhs.h_rhsDD[i][j]]
    pairs = sorted(zip(lhs_names, rhs_exprs), key=lambda p: p[0])
    evol_rhss = [lhrh(lhs=gri.gfaccess("rhs_gfs", n), rhs=e) for n, e in pairs]
    lp.LDS_STAGING = FD_ORDER // 2 + 1
    src, protos, costs = [], [], {}
    NE, NAE = table.n_evol, table.n_auxevol
    evol_fill = ("evol_gfs", 0, NE)

    # Ricci tensor (BaikalETK_C_kernels_codegen.py:306-330)
    ric_names = ["RbarDD%d%d" % (i, j) for i in range(3) for j in range(i, 3)]
    ric_exprs = {"RbarDD%d%d" % (i, j): RbarDD[i][j] for i in range(3) for j in range(i, 3)}
    kernels, flops = [], 0
    for k, grp in enumerate(ricci_groups(ric_names)):
        body = fin.FD_outputC("returnstring", [lhrh(lhs=gri.gfaccess("auxevol_gfs", n), rhs=ric_exprs[n])
                                               for n in grp], params="outCverbose=False")
        flops += bd.count_body(body)["flops"]
        name = "Baikal_BSSN_Ricci_g%d" % k
        text, launch = etk_kernel(name, "Ricci tensor components %s" % ", ".join(grp), body, table)
        src.append(text)
        kernels.append((name, launch))
    fname = "Baikal_BSSN_Ricci_FD_order_%d" % FD_ORDER
    c, p = entry_point(fname, "Ricci tensor RbarDD at interior points (BaikalETK_BSSN_Ricci_FD_order_4)",
                       ["const double *const *evol_gfs", "double *const *auxevol_gfs"],
                       [evol_fill, ("auxevol_gfs", 2 * NE, NAE)], kernels)
    src.append(c)
    protos.append(p)
    costs[fname] = flops
    print("(BENCH) Ricci codegen %.1fs" % (time.time() - t0), flush=True)

    # BSSN right-hand sides (BaikalETK_C_kernels_codegen.py:264-304)
    kernels, flops = [], 0
    for k, grp in enumerate(rhs_groups([n for n, _ in pairs])):
        if lds_tile_for_group(grp):
            lp.LDS_TILES["Baikal_BSSN_RHSs_g%d" % k] = lds_tile_for_group(grp)
        for outs, form in ETK_TILE_ADDRESS_OUTPUTS:
            if set(grp) == outs:
                addressing.TILE_ADDRESSING["Baikal_BSSN_RHSs_g%d" % k] = form
        if set(grp) == MIN_TWO_WAVES_OUTPUTS:
            lp.MIN_WAVES_PER_EU["Baikal_BSSN_RHSs_g%d" % k] = 2
        sel = [e for (n, _), e in zip(pairs, evol_rhss) if n in grp]
        body = fin.FD_outputC("returnstring", sel, params="outCverbose=False", upwindcontrolvec=betaU)
        flops += bd.count_body(body)["flops"]
        name = "Baikal_BSSN_RHSs_g%d" % k
        text, launch = etk_kernel(name, "BSSN right-hand sides of %s" % ", ".join(grp), body, table)
        src.append(text)
        kernels.append((name, launch))
    fname = "Baikal_BSSN_RHSs_enable_Tmunu_False_FD_order_%d" % FD_ORDER
    c, p = entry_point(fname, "BSSN right-hand sides at interior points, Ricci tensor read from RbarDD "
                       "(BaikalETK_BSSN_RHSs_enable_Tmunu_False_FD_order_4)",
                       ["const double *const *evol_gfs", "const double *const *auxevol_gfs",
                        "double *const *rhs_gfs"],
                       [evol_fill, ("auxevol_gfs", 2 * NE, NAE), ("rhs_gfs", NE, NE)], kernels)
    src.append(c)
    protos.append(p)
    costs[fname] = flops
    print("(BENCH) RHS codegen %.1fs" % (time.time() - t0), flush=True)

    # Hamiltonian and momentum constraints (BaikalETK_C_drivers_codegen.py:631-660), one kernel
    # each, as the SENR-layout bundle evaluates them
    kernels, flops = [], 0
    for kname, exprs in (("Baikal_BSSN_constraints_H", [lhrh(lhs=gri.gfaccess("aux_gfs", "H"), rhs=bssncon.H)]),
                         ("Baikal_BSSN_constraints_MU", [lhrh(lhs=gri.gfaccess("aux_gfs", "MU%d" % i),
                                                              rhs=bssncon.MU[i]) for i in range(3)])):
        body = fin.FD_outputC("returnstring", exprs, params="outCverbose=False")
        flops += bd.count_body(body)["flops"]
        text, launch = etk_kernel(kname, kname[len("Baikal_BSSN_constraints_"):] + " constraint", body, table)
        src.append(text)
        kernels.append((kname, launch))
    fname = "Baikal_BSSN_constraints_FD_order_%d" % FD_ORDER
    c, p = entry_point(fname, "H, MU0, MU1, MU2 at interior points (aux_gfs in that order: H, MU0, MU1, MU2)",
                       ["const double *const *evol_gfs", "double *const *aux_gfs"],
                       [evol_fill, ("aux_gfs", 2 * NE + NAE, table.n_aux)], kernels)
    src.append(c)
    protos.append(p)
    costs[fname] = flops

    # det(gammabar) = det(gammahat) at every point (BaikalETK_C_drivers_codegen.py:620-628)
    lp.LDS_STAGING = 0
    body = fin.FD_outputC("returnstring", detg_exprs, params="outCverbose=False")
    text, launch = etk_kernel("Baikal_enforce_detgammabar_constraint_k", "det(gammabar) enforcement", body, table,
                             interior=False)
    src.append(text)
    c, p = entry_point("Baikal_enforce_detgammabar_constraint",
                       "rescale hDD so det(gammabar) = det(gammahat), at every point",
                       ["double *const *evol_gfs"], [evol_fill], [("Baikal_enforce_detgammabar_constraint_k", launch)],
                       interior=False)
    src.append(c)
    protos.append(p)
    costs["Baikal_enforce_detgammabar_constraint"] = bd.count_body(body)["flops"]
    par.set_parval_from_str("grid::GridFuncMemAccess", "SENRlike")

    # Drivers that choose the kernel by FD_order (BaikalETK_C_drivers_codegen.py:533-585)
    drivers = ""
    for drv, target, args, call in (
            ("Baikal_driver_pt1_BSSN_Ricci", "Baikal_BSSN_Ricci_FD_order_%d",
             "const double *const *evol_gfs, double *const *auxevol_gfs", "evol_gfs, auxevol_gfs"),
            ("Baikal_driver_pt2_BSSN_RHSs", "Baikal_BSSN_RHSs_enable_Tmunu_False_FD_order_%d",
             "const double *const *evol_gfs, const double *const *auxevol_gfs, double *const *rhs_gfs",
             "evol_gfs, auxevol_gfs, rhs_gfs"),
            ("Baikal_BSSN_constraints", "Baikal_BSSN_constraints_FD_order_%d",
             "const double *const *evol_gfs, double *const *aux_gfs", "evol_gfs, aux_gfs")):
        sig = "const nrpy_cctkGH *cctkGH, const paramstruct *params, const int FD_order, %s, hipStream_t stream" % args
        drivers += ('extern "C" int %s(%s) {\n  if (FD_order == %d) return %s(cctkGH, params, %s, stream);\n'
                    "  return NRPY_ERR_PARAM;\n}\n" % (drv, sig, FD_ORDER, target % FD_ORDER, call))
        protos.append('extern "C" int %s(%s);' % (drv, sig))

    ps, members = bd.paramstruct_decl()
    too_wide = addressing.bound_check_code("(A0)", "(A1)") or "0"
    pre = ('#include "../nrpy_runtime_pre.h"\n#define NGHOSTS %d\n#define NRPY_ETK_NPTRS %d\n'
           % (FD_ORDER // 2 + 1, table.size) + ps + "\n" +
           "#define NRPY_ETK_TILE_TOO_WIDE(A0, A1) (%s)\n" % too_wide + GEOMETRY)
    meta = {"bundle": bundle_name, "family": "bssn_etk", "NGHOSTS": FD_ORDER // 2 + 1, "FD_order": FD_ORDER,
            "paramstruct": members,
            "gfs": {"evol": table.evol, "aux": table.aux, "auxevol": table.auxevol},
            "pointer_slots": {"evol": 0, "rhs": table.n_evol, "auxevol": 2 * table.n_evol,
                              "aux": 2 * table.n_evol + table.n_auxevol, "total": table.size},
            "lds_kernels": dict(lp.LDS_KERNELS), "cost": costs, "prototypes": protos,
            "tile_addressing": dict(sorted(addressing.APPLIED.items())),
            "kernels": list(costs)}
    return bd.write_bundle(bundle_name, [pre] + src + [drivers], meta)


if __name__ == "__main__":
    t = time.time()
    print("generated", bssn_etk(*sys.argv[1:2]), "in %.1fs" % (time.time() - t), flush=True)
