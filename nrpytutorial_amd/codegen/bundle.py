"""Assemble a system bundle: one HIP translation unit per evolution system.

A bundle is what one of the reference's Start-to-Finish builds generates for
one equation system and coordinate system, compiled into a shared library
with C linkage:

* ``paramstruct`` with every registered C parameter, members sorted by name
  (NRPy_param_funcs.py:227-243), gridfunction #defines in alphabetical order
  (grid.py:226-268), NGHOSTS;
* host functions ``set_Nxx_dxx_invdx_params__and__xx`` (reference_metric.py:1330-1395),
  ``xx_to_Cart`` (:1397-1409), ``find_timestep`` (:1412-1439), the EigenCoord
  maps and parity dot products used by the boundary builder
  (CurviBoundaryConditions.py:73-254);
* ``rfm_struct_create`` tabulating the precomputed reference-metric functions
  on the device (reference_metric.py:771-886);
* every kernel registered through ``outCfunction`` (rhs_eval, Ricci_eval, ...);
* fused kernels ``rhs_eval_rk4_substep`` that evaluate the RHS and apply the
  RK4 update of MoLtimestepping/C_Code_Generation.py:266-329 at each interior
  point, and ``MoL_step_forward_in_time`` / hipGraph replay that sequence
  [Ricci] -> RHS+update -> ghost fill -> [det(gammabar) enforcement] per substep
  as the generated RK_MoL.h does.
"""
import json
import os
import re
import sympy as sp

from .. import NRPy_param_funcs as par
from .. import grid as gri
from .. import reference_metric as rfm
from .. import outputC as oc
from .. import loop as lp

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
GENDIR = os.path.join(CSRC, "generated")


def _ctype(t):
    return {"REAL": "double", "int": "int", "bool": "int"}.get(t, "double")


def paramstruct_decl():
    members = par.paramstruct_members()
    s = "typedef struct __paramstruct__ {\n"
    for n, t in members:
        s += "  %s %s;\n" % (_ctype(t), n)
    s += "} paramstruct;\n"
    return s, [(n, _ctype(t), _default(n)) for n, t in members]


def _default(name):
    v = par.Cparam_value_default(name)
    try:
        return float(v)
    except (TypeError, ValueError):
        return 0.0


def _host_assign_block(exprs, names, used_params):
    """CSE'd C statements for host-side functions, with params unpacked."""
    code = oc.outputC(exprs, names, "returnstring", params="includebraces=False,outCverbose=False")
    return code


def _unpack_for(code, extra=()):
    idents = set(re.findall(r"\b[A-Za-z_][A-Za-z_0-9]*\b", code))
    out = ""
    for n, t in par.paramstruct_members():
        if n in idents and n not in ("xx0", "xx1", "xx2") and n not in extra:
            out += "  const %s %s = params->%s;\n" % (_ctype(t), n, n)
    return out


def coord_host_functions(CoordSystem):
    """Host C for xx_to_Cart, EigenCoord maps, parity dot products, xxmin/xxmax."""
    out = ""
    # xxmin/xxmax of the chosen and of the EigenCoord system
    par.set_parval_from_str("reference_metric::CoordSystem", CoordSystem)
    rfm.reference_metric(enable_compute_hatted_quantities=False)
    minmax = {}
    minmax[0] = ([oc.ccode(sp.sympify(v)) for v in rfm.xxmin], [oc.ccode(sp.sympify(v)) for v in rfm.xxmax])
    x2c = oc.outputC([rfm.xx_to_Cart[0], rfm.xx_to_Cart[1], rfm.xx_to_Cart[2]],
                     ["xCart[0]", "xCart[1]", "xCart[2]"], "returnstring",
                     params="includebraces=False,outCverbose=False")
    unit = rfm.UnitVectors
    xin = sp.symbols("xx0_inbounds xx1_inbounds xx2_inbounds", real=True)
    unit_in = [[sp.sympify(unit[i][j]).subs({rfm.xx[0]: xin[0], rfm.xx[1]: xin[1], rfm.xx[2]: xin[2]},
                                            simultaneous=True) for j in range(3)] for i in range(3)]
    parity = [sp.sympify(1)]
    for t in range(1, 4):
        parity.append(sum(unit[t - 1][i] * unit_in[t - 1][i] for i in range(3)))
    for i in range(3):
        for j in range(i, 3):
            parity.append(parity[i + 1] * parity[j + 1])
    pcode = oc.outputC(parity, ["parity[%d]" % i for i in range(10)], "returnstring",
                       params="includebraces=False,outCverbose=False")
    eigen = rfm.get_EigenCoord()
    par.set_parval_from_str("reference_metric::CoordSystem", eigen)
    rfm.reference_metric(enable_compute_hatted_quantities=False)
    minmax[1] = ([oc.ccode(sp.sympify(v)) for v in rfm.xxmin], [oc.ccode(sp.sympify(v)) for v in rfm.xxmax])
    ex2c = oc.outputC([rfm.xx_to_Cart[0], rfm.xx_to_Cart[1], rfm.xx_to_Cart[2]],
                      ["xCart[0]", "xCart[1]", "xCart[2]"], "returnstring",
                      params="includebraces=False,outCverbose=False")
    if any(sp.sympify(c) == 0 for c in rfm.Cart_to_xx[:3]):
        raise ValueError("EigenCoord " + eigen + " has no Cart_to_xx")
    c2x = oc.outputC([rfm.Cart_to_xx[0], rfm.Cart_to_xx[1], rfm.Cart_to_xx[2]],
                     ["Cart_to_xx[0]", "Cart_to_xx[1]", "Cart_to_xx[2]"], "returnstring",
                     params="includebraces=False,outCverbose=False")
    par.set_parval_from_str("reference_metric::CoordSystem", CoordSystem)
    rfm.reference_metric()

    def fn(sig, body, xxnames=True):
        pre = ""
        if xxnames:
            pre = "  const double xx0 = xx[0][i0];\n  const double xx1 = xx[1][i1];\n  const double xx2 = xx[2][i2];\n"
        return sig + " {\n" + _unpack_for(body) + pre + "".join("  " + l + "\n" for l in body.splitlines()) + "}\n"

    out += fn('extern "C" void xx_to_Cart(const paramstruct *params, double *xx[3], const int i0, const int i1, '
              'const int i2, double xCart[3])', x2c)
    out += fn("static void EigenCoord_xx_to_Cart(const paramstruct *params, double *xx[3], const int i0, "
              "const int i1, const int i2, double xCart[3])", ex2c)
    out += fn("static void EigenCoord_Cart_to_xx(const paramstruct *params, const double xCart[3], "
              "double Cart_to_xx[3])",
              "const double Cartx = xCart[0];\nconst double Carty = xCart[1];\nconst double Cartz = xCart[2];\n"
              + c2x, xxnames=False)
    out += fn("static void eval_parity_dot_products(const paramstruct *params, double parity[10], const double xx0, "
              "const double xx1, const double xx2, const double xx0_inbounds, const double xx1_inbounds, "
              "const double xx2_inbounds)", pcode, xxnames=False)
    # set_Nxx_dxx_invdx_params__and__xx
    body = ""
    for ec in (0, 1):
        mn, mx = minmax[ec]
        body += "  if (EigenCoord == %d) {\n" % ec
        for d in range(3):
            body += "    xxmin[%d] = %s;\n    xxmax[%d] = %s;\n" % (d, mn[d], d, mx[d])
        body += "  }\n"
    allcode = body
    unpack = ""
    idents = set(re.findall(r"\b[A-Za-z_][A-Za-z_0-9]*\b", allcode))
    for n, t in par.paramstruct_members():
        if n in idents:
            unpack += "  const %s %s = params->%s;\n" % (_ctype(t), n, n)
    out += ('extern "C" int set_Nxx_dxx_invdx_params__and__xx(const int EigenCoord, const int Nxx[3], '
            'paramstruct *params, double *xx[3]) {\n'
            "  if (!params || !Nxx || !xx) return NRPY_ERR_NULL;\n"
            "  if (EigenCoord != 0 && EigenCoord != 1) return NRPY_ERR_PARAM;\n"
            "  for (int d = 0; d < 3; d++) if (Nxx[d] < 1) return NRPY_ERR_SHAPE;\n"
            "  params->Nxx0 = Nxx[0]; params->Nxx1 = Nxx[1]; params->Nxx2 = Nxx[2];\n"
            "  params->Nxx_plus_2NGHOSTS0 = Nxx[0] + 2*NGHOSTS;\n"
            "  params->Nxx_plus_2NGHOSTS1 = Nxx[1] + 2*NGHOSTS;\n"
            "  params->Nxx_plus_2NGHOSTS2 = Nxx[2] + 2*NGHOSTS;\n"
            + unpack +
            "  double xxmin[3], xxmax[3];\n" + body +
            "  params->dxx0 = (xxmax[0] - xxmin[0]) / ((double)Nxx[0]);\n"
            "  params->dxx1 = (xxmax[1] - xxmin[1]) / ((double)Nxx[1]);\n"
            "  params->dxx2 = (xxmax[2] - xxmin[2]) / ((double)Nxx[2]);\n"
            "  params->invdx0 = 1.0/params->dxx0;\n"
            "  params->invdx1 = 1.0/params->dxx1;\n"
            "  params->invdx2 = 1.0/params->dxx2;\n"
            "  const int NN[3] = {params->Nxx_plus_2NGHOSTS0, params->Nxx_plus_2NGHOSTS1, params->Nxx_plus_2NGHOSTS2};\n"
            "  const double dd[3] = {params->dxx0, params->dxx1, params->dxx2};\n"
            "  for (int d = 0; d < 3; d++) {\n"
            "    if (!xx[d]) return NRPY_ERR_NULL;\n"
            "    for (int j = 0; j < NN[d]; j++) xx[d][j] = xxmin[d] + ((double)(j - NGHOSTS) + (1.0/2.0))*dd[d];\n"
            "  }\n"
            "  return NRPY_OK;\n}\n")
    # find_timestep: dt = CFL * min over interior of min_d (scale factor_d * dxx_d) / wavespeed
    dx = sp.symbols("dxx0 dxx1 dxx2", real=True)
    ds = oc.outputC([sp.sympify(rfm.scalefactor_orthog[d]) * dx[d] for d in range(3)],
                    ["ds_dirn0", "ds_dirn1", "ds_dirn2"], "returnstring",
                    params="includebraces=False,outCverbose=False")
    out += ('extern "C" double find_timestep(const paramstruct *params, double *xx[3], const double CFL_FACTOR) {\n'
            + _unpack_for(ds) +
            "  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;\n"
            "  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;\n"
            "  const int Nxx_plus_2NGHOSTS2 = params->Nxx_plus_2NGHOSTS2;\n"
            "  double dsmin = 1e38;\n"
            "  for (int i2 = NGHOSTS; i2 < Nxx_plus_2NGHOSTS2 - NGHOSTS; i2++)\n"
            "  for (int i1 = NGHOSTS; i1 < Nxx_plus_2NGHOSTS1 - NGHOSTS; i1++)\n"
            "  for (int i0 = NGHOSTS; i0 < Nxx_plus_2NGHOSTS0 - NGHOSTS; i0++) {\n"
            "    const double xx0 = xx[0][i0]; const double xx1 = xx[1][i1]; const double xx2 = xx[2][i2];\n"
            "    (void)xx0; (void)xx1; (void)xx2;\n"
            "    double ds_dirn0, ds_dirn1, ds_dirn2;\n"
            + "".join("    " + l.replace("const double ds_dirn", "ds_dirn") + "\n" for l in ds.splitlines()) +
            "    dsmin = fmin(dsmin, fmin(ds_dirn0, fmin(ds_dirn1, ds_dirn2)));\n"
            "  }\n"
            "  return dsmin*CFL_FACTOR/%s;\n}\n" % _wavespeed_expr())
    return out


def _wavespeed_expr():
    """find_timestep's divisor (reference_metric.py:1434): the wavespeed C parameter where the
    system registers one (the scalar wave), else 1.0, the value the BSSN and Maxwell
    Start-to-Finish drivers #define."""
    return "params->wavespeed" if any(n == "wavespeed" for n, _ in par.paramstruct_members()) else "1.0"


def rfm_struct_code():
    """rfm_struct (device table pointers) and its host-side constructor."""
    tables = rfm.rfm_precompute_tables()
    s = "typedef struct __rfmstruct__ {\n"
    if not tables:
        s += "  const double *unused;\n"
    for name, expr, axes in tables:
        s += "  const double *%s;\n" % name
    s += "} rfm_struct;\n"
    body = ""
    for k, (name, expr, axes) in enumerate(tables):
        code = oc.outputC(expr, "v", "returnstring", params="includebraces=False,outCverbose=False")
        unpack = _unpack_for(code)
        if len(axes) == 1:
            d = axes[0]
            body += ("  { std::vector<double> h(NN[%d]);\n    for (int i%d = 0; i%d < NN[%d]; i%d++) {\n"
                     "      const double xx%d = xx[%d][i%d];\n      double v;\n%s%s      h[i%d] = v;\n    }\n"
                     "    if (!nrpy_upload(&tab[%d], h.data(), h.size())) rc = NRPY_ERR_ALLOC; }\n" %
                     (d, d, d, d, d, d, d, d, "".join("    " + l + "\n" for l in unpack.splitlines()),
                      "".join("      " + l.replace("const double v", "v") + "\n" for l in code.splitlines()),
                      d, k))
        else:
            body += ("  { std::vector<double> h((size_t)NN[0]*NN[1]);\n"
                     "    for (int i1 = 0; i1 < NN[1]; i1++) for (int i0 = 0; i0 < NN[0]; i0++) {\n"
                     "      const double xx0 = xx[0][i0]; const double xx1 = xx[1][i1];\n      double v;\n%s%s"
                     "      h[i0 + NN[0]*i1] = v;\n    }\n"
                     "    if (!nrpy_upload(&tab[%d], h.data(), h.size())) rc = NRPY_ERR_ALLOC; }\n" %
                     ("".join("    " + l + "\n" for l in unpack.splitlines()),
                      "".join("      " + l.replace("const double v", "v") + "\n" for l in code.splitlines()), k))
    n = max(1, len(tables))
    assign = "".join("  r->%s = tab[%d];\n" % (name, k) for k, (name, _, _) in enumerate(tables))
    frees = "".join("  (void)hipFree((void *)rfmstruct->%s);\n" % name for name, _, _ in tables)
    s += ('extern "C" int rfm_struct_create(const paramstruct *params, double *xx[3], rfm_struct **out) {\n'
          "  if (!params || !xx || !out) return NRPY_ERR_NULL;\n"
          "  *out = nullptr;\n"
          "  const int NN[3] = {params->Nxx_plus_2NGHOSTS0, params->Nxx_plus_2NGHOSTS1, params->Nxx_plus_2NGHOSTS2};\n"
          "  (void)NN;\n"
          "  double *tab[%d] = {0};\n  int rc = NRPY_OK;\n%s"
          "  rfm_struct *r = new rfm_struct();\n  memset(r, 0, sizeof(*r));\n%s"
          "  *out = r;\n  return rc;\n}\n"
          'extern "C" void rfm_struct_free(rfm_struct *rfmstruct) {\n  if (!rfmstruct) return;\n%s'
          "  delete rfmstruct;\n}\n" % (n, body, assign, frees))
    return s


def check_params_code():
    from . import lds, addressing
    # kernels whose x-neighbours come from lane moves need rows of at least 2 HX points
    nx_min = max(1, lds.MIN_NXX0)
    # kernels with tile-relative addressing need their 32-bit lane offsets below 2^32 bytes
    tile = addressing.bound_check_code()
    return ("static int nrpy_check_params(const paramstruct *p) {\n"
            "  if (p->Nxx0 < %d || p->Nxx1 < 1 || p->Nxx2 < 1) return 1;\n" % nx_min +
            "  if (p->Nxx_plus_2NGHOSTS0 != p->Nxx0 + 2*NGHOSTS || p->Nxx_plus_2NGHOSTS1 != p->Nxx1 + 2*NGHOSTS ||\n"
            "      p->Nxx_plus_2NGHOSTS2 != p->Nxx2 + 2*NGHOSTS) return 1;\n"
            "  if ((double)p->Nxx_plus_2NGHOSTS0 * p->Nxx_plus_2NGHOSTS1 * p->Nxx_plus_2NGHOSTS2 > 2147483647.0) return 1;\n" +
            ("  if (%s) return 1;\n" % tile if tile else "") +
            "  return 0;\n}\n")


WRITE_RE = re.compile(r"rhs_gfs\[IDX4\((\w+)GF, i0, i1, i2\)\] = (.*);")


def fused_rk4_kernel(name, args_decl, call_args, body_with_writes, loopopts="InteriorPoints,Enable_rfm_precompute"):
    """Template kernel evaluating the RHS and applying RK4 substep S at each interior point."""
    # The RHS values are computed first; the RK update (loads of y_n and the
    # running total) comes after a scheduling barrier so the compiler does not
    # hoist those loads to the top of the kernel, where they would stay live
    # through the whole RHS evaluation.
    lines = body_with_writes.splitlines()
    vals = []
    for i, l in enumerate(lines):
        m = WRITE_RE.search(l)
        if m:
            lines[i] = "const double __rhs_%s = %s;" % (m.group(1), m.group(2))
            vals.append(m.group(1))
    tail = (["__builtin_amdgcn_sched_barrier(0);"] +
            ["NRPY_RK4_PRELOAD(S, %sGF, %d)" % (v, k) for k, v in enumerate(vals)] +
            ["NRPY_RK4_STORE(S, %d, __rhs_%s);" % (k, v) for k, v in enumerate(vals)])
    # insert inside the body's closing brace, where the __rhs_* values are in scope
    close = max(i for i, l in enumerate(lines) if l.strip() == "}")
    lines[close:close] = tail
    fused_body = "\n".join(lines)
    if "rhs_gfs" in fused_body:
        raise ValueError("unrecognised write to rhs_gfs in the RHS body")
    kbody = lp.simple_loop(loopopts, fused_body)
    arrays = [("const double *" if "const" in a else "double *", a.split()[-1])
              for a in args_decl.split(",") if "*" in a]
    # rk_yn and rk_out are the same array in the last substep: no __restrict__ on the RK arrays.
    lp.BATCH_SPECS[name] = {"arrays": [(q + "__restrict__ ", an) for q, an in arrays] +
                                      [("double *", "rk_yn"), ("double *", "rk_rt"), ("double *", "rk_out")],
                            "rfm": "rfm_struct" in args_decl, "template": "S", "body": fused_body,
                            "loopopts": loopopts, "dt": True}
    unpack = "".join("  const %s %s = params.%s;\n" % (lp._ctype(t), n, n) for n, t in lp._used_cparams(kbody))
    for n in ("Nxx0", "Nxx1", "Nxx2", "Nxx_plus_2NGHOSTS0", "Nxx_plus_2NGHOSTS1", "Nxx_plus_2NGHOSTS2"):
        if not re.search(r"const int %s = " % n, unpack):
            unpack += "  const int %s = params.%s;\n" % (n, n)
    kernel = ("template <int S>\n__global__ void __launch_bounds__(%s) %s__gpu(%s, double *rk_yn, double *rk_rt, "
              "double *rk_out, const double dt) {\n%s%s}\n" %
              (lp.launch_bounds(name), name, args_decl, unpack, "".join("  " + l + "\n" for l in kbody.splitlines())))
    if "InteriorPoints" in loopopts:
        kernel = lp.stage_lds(name, kernel)[0]
    return kernel


_TOKEN = re.compile(r"[A-Za-z_][A-Za-z_0-9]*|\d+\.\d*(?:[eE][-+]?\d+)?|\d+|\S")
_FUNCS = ("sqrt", "cbrt", "pow", "exp", "log", "sin", "cos", "tan", "fabs", "sinh", "cosh", "tanh", "asinh",
          "acos", "atan2")


def count_body(body):
    """Algorithmic cost of one gridpoint of a generated body.

    FLOPs: every binary +, -, *, / and every math-library call counts 1, so a
    fused multiply-add counts 2. Bytes: 8 per distinct (array, gridfunction)
    read and per distinct gridfunction written, i.e. each input read once and
    each output written once (stencil re-reads are not counted)."""
    flops = 0
    reads, writes = set(), set()
    for line in body.splitlines():
        line = line.strip()
        if "=" not in line or line.startswith("/*") or line.startswith("//"):
            continue
        lhs, rhs = line.split("=", 1)
        m = re.search(r"(\w+)\[IDX4\((\w+)GF", lhs) or re.search(r"()(\w+)GF\[CCTK_GFINDEX3D\(", lhs)
        if m:
            writes.add((m.group(1), m.group(2)))
        for a, g in re.findall(r"(\w+)\[IDX4\((\w+)GF", rhs) + re.findall(r"()(\w+)GF\[CCTK_GFINDEX3D\(", rhs):
            reads.add((a, g))
        if "[IDX4(" in rhs and re.match(r"\s*\w+\[IDX4\([^)]*\)\];\s*$", rhs):
            continue  # plain load
        if re.match(r"\s*\w+GF\[CCTK_GFINDEX3D\([^\]]*\)\];\s*$", rhs):
            continue  # plain load, Einstein Toolkit form
        toks = _TOKEN.findall(rhs.rstrip(";"))
        prev_operand = False
        for t in toks:
            if t in ("+", "-"):
                if prev_operand:
                    flops += 1
                prev_operand = False
            elif t in ("*", "/"):
                flops += 1
                prev_operand = False
            elif t in _FUNCS:
                flops += 1
                prev_operand = False
            elif t == "(" or t == ",":
                prev_operand = False
            elif t in ("?", ":", ">", "<"):
                prev_operand = False
            else:
                prev_operand = True
    return {"flops": flops, "read_gfs": len(reads), "write_gfs": len(writes),
            "bytes": 8 * (len(reads) + len(writes))}


def write_bundle(bundle_name, parts, meta):
    """Write csrc/generated/<bundle>.hip and <bundle>.json (or into $NRPY_GENDIR, a sibling
    directory of csrc/generated, when generating variants for tuning)."""
    gendir = os.environ.get("NRPY_GENDIR", GENDIR)
    os.makedirs(gendir, exist_ok=True)
    from .. import finite_difference as fin
    fdfuncs = fin.finite_difference_functions_code()
    if fdfuncs:
        # operator functions of FD_functions_enable go after the header (first part), before any kernel
        parts = [parts[0], fdfuncs] + list(parts[1:])
    src = "// Generated by nrpytutorial_amd.codegen; do not edit.\n" + "\n".join(parts)
    with open(os.path.join(gendir, bundle_name + ".hip"), "w") as f:
        f.write(src)
    with open(os.path.join(gendir, bundle_name + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    return os.path.join(gendir, bundle_name + ".hip")


def gf_defines_and_meta():
    evol, aux, auxevol = gri.gridfunction_lists()
    return gri.gridfunction_defines(), {"evol": evol, "aux": aux, "auxevol": auxevol}


def parity_type(name):
    """Parity type of a gridfunction from its rank and trailing digits (CurviBoundaryConditions.py:122-185)."""
    gf = [g for g in gri.glb_gridfcs_list if g.name == name][0]
    if gf.rank == 0:
        return 0
    if gf.rank == 1:
        return int(name[-1]) + 1
    a, b = sorted((int(name[-2]), int(name[-1])))
    return {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}[(a, b)]
