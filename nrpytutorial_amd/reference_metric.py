"""Reference metric: coordinate systems, hatted quantities and rfm precompute.

Follows reference_metric.py:

* coordinate systems Spherical / SinhSpherical / SinhSphericalv2
  (reference_metric.py:100-179), Cylindrical / SinhCylindrical /
  SinhCylindricalv2 (:226-304) and Cartesian (:398-428): xxmin/xxmax,
  xxSph, xx_to_Cart, Cart_to_xx, orthogonal scale factors, unit vectors;
* ``ref_metric__hatted_quantities`` builds ghat_ij = diag(s_i^2), the
  rescaling vector ReU = 1/s_i and matrix ReDD = s_i s_j, their first and
  second derivatives and the hatted Christoffel symbols (:487-601);
* with ``reference_metric::enable_rfm_precompute`` the scale factors are
  written through generic functions f0_of_xx0, f1_of_xx1, f2_of_xx0_xx1,
  f3_of_xx0, f4_of_xx2 whose derivatives become symbols such as
  ``f1_of_xx1__DD11`` (:605-690); derivatives that are constants are folded
  in (:692-770), and the remaining symbols are tabulated on the grid
  (``rfm_precompute_tables``) instead of being evaluated at every point.
"""
import sympy as sp
from . import NRPy_param_funcs as par
from . import grid as gri
from . import indexedexp as ixp

thismodule = "reference_metric"
par.initialize_param(par.glb_param("char", thismodule, "CoordSystem", "Spherical"))
par.initialize_param(par.glb_param("char", thismodule, "enable_rfm_precompute", "False"))
par.initialize_param(par.glb_param("char", thismodule, "rfm_precompute_Ccode_outdir", "Ccode"))

xx = gri.xx
Cartx, Carty, Cartz = sp.symbols("Cartx Carty Cartz", real=True)
Cart = [Cartx, Carty, Cartz]

xx_to_Cart = ixp.zerorank1(DIM=4)
Cart_to_xx = ixp.zerorank1(DIM=4)
xxSph = ixp.zerorank1(DIM=4)
scalefactor_orthog = ixp.zerorank1(DIM=4)
scalefactor_orthog_funcform = ixp.zerorank1(DIM=4)
UnitVectors = ixp.zerorank2(DIM=3)
xxmin = [0, 0, 0]
xxmax = [0, 0, 0]
have_already_called_reference_metric_function = False

# Generic-function names used by rfm precompute, and the closed forms the
# chosen coordinate system assigns to them.
_FUNC_NAMES = ("f0_of_xx0", "f1_of_xx1", "f2_of_xx0_xx1", "f3_of_xx0", "f4_of_xx2")
_funcform = {}
_closed_form = {}
# symbol name -> closed-form expression, for the precomputed tables
rfm_precompute_symbols = {}


def nrpysinh(x):
    return (sp.exp(x) - sp.exp(-x)) * sp.Rational(1, 2)


def nrpycosh(x):
    return (sp.exp(x) + sp.exp(-x)) * sp.Rational(1, 2)


def _reset_globals():
    global xx_to_Cart, Cart_to_xx, xxSph, scalefactor_orthog, scalefactor_orthog_funcform, UnitVectors
    xx_to_Cart = ixp.zerorank1(DIM=4)
    Cart_to_xx = ixp.zerorank1(DIM=4)
    xxSph = ixp.zerorank1(DIM=4)
    scalefactor_orthog = ixp.zerorank1(DIM=4)
    scalefactor_orthog_funcform = ixp.zerorank1(DIM=4)
    UnitVectors = ixp.zerorank2(DIM=3)


def reference_metric(SymPySimplifyExpressions=True, enable_compute_hatted_quantities=True):
    """Set up the coordinate system named by ``reference_metric::CoordSystem``."""
    global xxmin, xxmax, UnitVectors, have_already_called_reference_metric_function
    _reset_globals()
    have_already_called_reference_metric_function = True
    xx[0], xx[1], xx[2] = sp.symbols("xx0 xx1 xx2", real=True)
    _funcform.clear()
    _funcform["f0_of_xx0"] = sp.Function("f0_of_xx0_funcform")(xx[0])
    _funcform["f1_of_xx1"] = sp.Function("f1_of_xx1_funcform")(xx[1])
    _funcform["f2_of_xx0_xx1"] = sp.Function("f2_of_xx0_xx1_funcform")(xx[0], xx[1])
    _funcform["f3_of_xx0"] = sp.Function("f3_of_xx0_funcform")(xx[0])
    _funcform["f4_of_xx2"] = sp.Function("f4_of_xx2_funcform")(xx[2])
    _closed_form.clear()

    CoordSystem = par.parval_from_str("reference_metric::CoordSystem")
    M_PI, M_SQRT1_2 = par.Cparameters("#define", thismodule, ["M_PI", "M_SQRT1_2"], "")

    if CoordSystem in ("Spherical", "SinhSpherical", "SinhSphericalv2"):
        if CoordSystem == "Spherical":
            RMAX = par.Cparameters("REAL", thismodule, ["RMAX"], 10.0)
            xxmin = [sp.sympify(0), sp.sympify(0), -M_PI]
            xxmax = [RMAX, M_PI, M_PI]
            r = xx[0]
            Cart_to_xx[0] = sp.sqrt(Cartx ** 2 + Carty ** 2 + Cartz ** 2)
            Cart_to_xx[1] = sp.acos(Cartz / Cart_to_xx[0])
            Cart_to_xx[2] = sp.atan2(Carty, Cartx)
        else:
            xxmin = [sp.sympify(0), sp.sympify(0), -M_PI]
            xxmax = [sp.sympify(1), M_PI, M_PI]
            AMPL, SINHW = par.Cparameters("REAL", thismodule, ["AMPL", "SINHW"], [10.0, 0.2])
            sinh_part = (sp.exp(xx[0] / SINHW) - sp.exp(-xx[0] / SINHW)) / (sp.exp(1 / SINHW) - sp.exp(-1 / SINHW))
            if CoordSystem == "SinhSpherical":
                r = AMPL * sinh_part
                Cart_to_xx[0] = SINHW * sp.asinh(sp.sqrt(Cartx ** 2 + Carty ** 2 + Cartz ** 2) * sp.sinh(1 / SINHW) / AMPL)
                Cart_to_xx[1] = sp.acos(Cartz / sp.sqrt(Cartx ** 2 + Carty ** 2 + Cartz ** 2))
                Cart_to_xx[2] = sp.atan2(Carty, Cartx)
            else:
                const_dr = par.Cparameters("REAL", thismodule, ["const_dr"], 0.0625)
                r = AMPL * (const_dr * xx[0] + sinh_part)
                # no closed-form inverse of the radial map
        th, ph = xx[1], xx[2]
        xxSph[0], xxSph[1], xxSph[2] = r, th, ph
        xx_to_Cart[0] = r * sp.sin(th) * sp.cos(ph)
        xx_to_Cart[1] = r * sp.sin(th) * sp.sin(ph)
        xx_to_Cart[2] = r * sp.cos(th)
        scalefactor_orthog[0] = sp.diff(r, xx[0])
        scalefactor_orthog[1] = r
        scalefactor_orthog[2] = r * sp.sin(th)
        _closed_form["f0_of_xx0"] = r
        _closed_form["f1_of_xx1"] = sp.sin(th)
        f0, f1 = _funcform["f0_of_xx0"], _funcform["f1_of_xx1"]
        scalefactor_orthog_funcform[0] = sp.diff(f0, xx[0])
        scalefactor_orthog_funcform[1] = f0
        scalefactor_orthog_funcform[2] = f0 * f1
        UnitVectors = [[sp.sin(th) * sp.cos(ph), sp.sin(th) * sp.sin(ph), sp.cos(th)],
                       [sp.cos(th) * sp.cos(ph), sp.cos(th) * sp.sin(ph), -sp.sin(th)],
                       [-sp.sin(ph), sp.cos(ph), sp.sympify(0)]]

    elif CoordSystem in ("Cylindrical", "SinhCylindrical", "SinhCylindricalv2"):
        if CoordSystem == "Cylindrical":
            RHOMAX, ZMIN, ZMAX = par.Cparameters("REAL", thismodule, ["RHOMAX", "ZMIN", "ZMAX"], [10.0, -10.0, 10.0])
            xxmin = [sp.sympify(0), -M_PI, ZMIN]
            xxmax = [RHOMAX, M_PI, ZMAX]
            RHOCYL, PHICYL, ZCYL = xx[0], xx[1], xx[2]
            Cart_to_xx[0] = sp.sqrt(Cartx ** 2 + Carty ** 2)
            Cart_to_xx[1] = sp.atan2(Carty, Cartx)
            Cart_to_xx[2] = Cartz
        else:
            xxmin = [sp.sympify(0), -M_PI, sp.sympify(-1)]
            xxmax = [sp.sympify(1), M_PI, sp.sympify(+1)]
            AMPLRHO, SINHWRHO, AMPLZ, SINHWZ = par.Cparameters(
                "REAL", thismodule, ["AMPLRHO", "SINHWRHO", "AMPLZ", "SINHWZ"], [10.0, 0.2, 10.0, 0.2])
            srho = (sp.exp(xx[0] / SINHWRHO) - sp.exp(-xx[0] / SINHWRHO)) / (sp.exp(1 / SINHWRHO) - sp.exp(-1 / SINHWRHO))
            sz = (sp.exp(xx[2] / SINHWZ) - sp.exp(-xx[2] / SINHWZ)) / (sp.exp(1 / SINHWZ) - sp.exp(-1 / SINHWZ))
            PHICYL = xx[1]
            if CoordSystem == "SinhCylindrical":
                RHOCYL = AMPLRHO * srho
                ZCYL = AMPLZ * sz
                Cart_to_xx[0] = SINHWRHO * sp.asinh(sp.sqrt(Cartx ** 2 + Carty ** 2) * sp.sinh(1 / SINHWRHO) / AMPLRHO)
                Cart_to_xx[1] = sp.atan2(Carty, Cartx)
                Cart_to_xx[2] = SINHWZ * sp.asinh(Cartz * sp.sinh(1 / SINHWZ) / AMPLZ)
            else:
                const_drho, const_dz = par.Cparameters("REAL", thismodule, ["const_drho", "const_dz"], [0.0625, 0.0625])
                RHOCYL = AMPLRHO * (const_drho * xx[0] + srho)
                ZCYL = AMPLZ * (const_dz * xx[2] + sz)
        xx_to_Cart[0] = RHOCYL * sp.cos(PHICYL)
        xx_to_Cart[1] = RHOCYL * sp.sin(PHICYL)
        xx_to_Cart[2] = ZCYL
        xxSph[0] = sp.sqrt(RHOCYL ** 2 + ZCYL ** 2)
        xxSph[1] = sp.acos(ZCYL / xxSph[0])
        xxSph[2] = PHICYL
        scalefactor_orthog[0] = sp.diff(RHOCYL, xx[0])
        scalefactor_orthog[1] = RHOCYL
        scalefactor_orthog[2] = sp.diff(ZCYL, xx[2])
        _closed_form["f0_of_xx0"] = RHOCYL
        _closed_form["f4_of_xx2"] = sp.diff(ZCYL, xx[2])
        f0, f4 = _funcform["f0_of_xx0"], _funcform["f4_of_xx2"]
        scalefactor_orthog_funcform[0] = sp.diff(f0, xx[0])
        scalefactor_orthog_funcform[1] = f0
        scalefactor_orthog_funcform[2] = f4
        UnitVectors = [[sp.cos(PHICYL), sp.sin(PHICYL), sp.sympify(0)],
                       [-sp.sin(PHICYL), sp.cos(PHICYL), sp.sympify(0)],
                       [sp.sympify(0), sp.sympify(0), sp.sympify(1)]]

    elif CoordSystem in ("SymTP", "SinhSymTP"):
        # Prolate-spheroidal ("symmetric two-puncture") coordinates (reference_metric.py:306-393):
        # rho = A sin(xx1), z = sqrt(A^2 + b^2) cos(xx1), phi = xx2, with A = xx0 (SymTP) or a
        # sinh map of xx0 to [0, AMAX] (SinhSymTP). The foci sit at z = +-bScale.
        bScale, SINHWAA, AMAX = par.Cparameters("REAL", thismodule, ["bScale", "SINHWAA", "AMAX"], [0.5, 0.2, 10.0])
        xxmin = [sp.sympify(0), sp.sympify(0), -M_PI]
        xxmax = [AMAX, M_PI, M_PI]
        AA = xx[0]
        if CoordSystem == "SinhSymTP":
            xxmax[0] = sp.sympify(1)
            AA = AMAX * (sp.exp(xx[0] / SINHWAA) - sp.exp(-xx[0] / SINHWAA)) / (sp.exp(1 / SINHWAA) - sp.exp(-1 / SINHWAA))
        th, ph = xx[1], xx[2]
        var1 = sp.sqrt(AA ** 2 + (bScale * sp.sin(th)) ** 2)
        var2 = sp.sqrt(AA ** 2 + bScale ** 2)
        rho, z = AA * sp.sin(th), var2 * sp.cos(th)
        xx_to_Cart[0] = AA * sp.sin(th) * sp.cos(ph)
        xx_to_Cart[1] = AA * sp.sin(th) * sp.sin(ph)
        xx_to_Cart[2] = z
        xxSph[0] = sp.sqrt(rho ** 2 + z ** 2)
        xxSph[1] = sp.acos(z / xxSph[0])
        xxSph[2] = ph
        if CoordSystem == "SymTP":
            # inverse map: A^2 and cos^2(xx1) are the roots of a quadratic in terms of the
            # spherical radius and polar angle of the Cartesian point
            rS = sp.sqrt(Cartx ** 2 + Carty ** 2 + Cartz ** 2)
            thS = sp.acos(Cartz / rS)
            disc = sp.sqrt(bScale ** 4 + 2 * bScale ** 2 * rS ** 2 + rS ** 4 - 4 * bScale ** 2 * rS ** 2 * sp.cos(thS) ** 2)
            Cart_to_xx[0] = sp.sqrt(-bScale ** 2 + rS ** 2 + disc) * M_SQRT1_2
            Cart_to_xx[1] = sp.acos(sp.sign(Cartz) * (sp.sqrt(1 + rS ** 2 / bScale ** 2 - disc / bScale ** 2) * M_SQRT1_2))
            Cart_to_xx[2] = sp.atan2(Carty, Cartx)
        scalefactor_orthog[0] = sp.diff(AA, xx[0]) * var1 / var2
        scalefactor_orthog[1] = var1
        scalefactor_orthog[2] = AA * sp.sin(th)
        _closed_form["f0_of_xx0"] = AA
        _closed_form["f1_of_xx1"] = sp.sin(th)
        _closed_form["f2_of_xx0_xx1"] = var1
        _closed_form["f3_of_xx0"] = var2
        f0, f1, f2, f3 = (_funcform[k] for k in ("f0_of_xx0", "f1_of_xx1", "f2_of_xx0_xx1", "f3_of_xx0"))
        scalefactor_orthog_funcform[0] = sp.diff(f0, xx[0]) * f2 / f3
        scalefactor_orthog_funcform[1] = f2
        scalefactor_orthog_funcform[2] = f0 * f1
        UnitVectors = [[sp.sin(th) * sp.cos(ph) * var2 / var1, sp.sin(th) * sp.sin(ph) * var2 / var1,
                        AA * sp.cos(th) / var1],
                       [AA * sp.cos(th) * sp.cos(ph) / var1, AA * sp.cos(th) * sp.sin(ph) / var1,
                        -sp.sin(th) * var2 / var1],
                       [-sp.sin(ph), sp.cos(ph), sp.sympify(0)]]

    elif CoordSystem == "Cartesian":
        par.Cparameters("REAL", thismodule, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"],
                        [-10.0, 10.0, -10.0, 10.0, -10.0, 10.0])
        xxmin = [sp.Symbol("xmin", real=True), sp.Symbol("ymin", real=True), sp.Symbol("zmin", real=True)]
        xxmax = [sp.Symbol("xmax", real=True), sp.Symbol("ymax", real=True), sp.Symbol("zmax", real=True)]
        for i in range(3):
            xx_to_Cart[i] = xx[i]
            Cart_to_xx[i] = Cart[i]
            scalefactor_orthog[i] = sp.sympify(1)
            scalefactor_orthog_funcform[i] = sp.sympify(1)
        xxSph[0] = sp.sqrt(xx[0] ** 2 + xx[1] ** 2 + xx[2] ** 2)
        xxSph[1] = sp.acos(xx[2] / xxSph[0])
        xxSph[2] = sp.atan2(xx[1], xx[0])
        UnitVectors = [[sp.sympify(int(i == j)) for j in range(3)] for i in range(3)]
    else:
        raise ValueError("CoordSystem == " + CoordSystem + " is not supported")

    if enable_compute_hatted_quantities:
        ref_metric__hatted_quantities(SymPySimplifyExpressions)


def _derivative_symbol_name(d):
    """Derivative(f0_of_xx0_funcform(xx0), (xx0, 2)) -> 'f0_of_xx0__DD00'."""
    base = str(d.args[0].func).replace("_funcform", "")
    suffix = ""
    for var, count in d.variable_count:
        suffix += str(var)[2:] * int(count)
    return base + "__" + "D" * len(suffix) + suffix


def _to_precompute_symbols(expr):
    expr = sp.sympify(expr)
    rule = {}
    for node in sp.preorder_traversal(expr):
        if isinstance(node, sp.Derivative):
            rule[node] = sp.Symbol(_derivative_symbol_name(node), real=True)
    expr = expr.xreplace(rule)
    rule = {}
    for node in sp.preorder_traversal(expr):
        if isinstance(node, sp.core.function.AppliedUndef) and "_funcform" in str(node.func):
            rule[node] = sp.Symbol(str(node.func).replace("_funcform", ""), real=True)
    return expr.xreplace(rule)


def _closed_form_of_precompute_symbol(name):
    base, _, deriv = name.partition("__")
    f = _closed_form.get(base)
    if f is None:
        raise ValueError("reference metric function " + base + " is not defined for this CoordSystem")
    for c in deriv.replace("D", ""):
        f = sp.diff(f, xx[int(c)])
    return f


def _map_tensor(fn, obj):
    if isinstance(obj, list):
        return [_map_tensor(fn, o) for o in obj]
    return fn(obj)


def ref_metric__hatted_quantities(SymPySimplifyExpressions=True):
    """Hatted metric, rescaling quantities, their derivatives, Christoffels
    (reference_metric.py:487-601) and, with precompute on, the substitution of
    generic functions (reference_metric.py:605-770)."""
    global ReU, ReDD, ghatDD, ghatUU, detgammahat, detgammahatdD, detgammahatdDD
    global ReUdD, ReUdDD, ReDDdD, ReDDdDD, ghatDDdD, ghatDDdDD, GammahatUDD, GammahatUDDdD
    precompute = par.parval_from_str("reference_metric::enable_rfm_precompute") == "True"
    DIM = 3
    sf = [sp.sympify(scalefactor_orthog_funcform[i] if precompute else scalefactor_orthog[i]) for i in range(DIM)]
    ReU = [1 / sf[i] for i in range(DIM)]
    ReDD = [[sf[i] * sf[j] for j in range(DIM)] for i in range(DIM)]
    ghatDD = ixp.zerorank2()
    for i in range(DIM):
        ghatDD[i][i] = sf[i] ** 2
    ghatUU, detgammahat = ixp.symm_matrix_inverter3x3(ghatDD)

    detgammahatdD = [sp.diff(detgammahat, xx[i]) for i in range(DIM)]
    detgammahatdDD = [[sp.diff(detgammahatdD[i], xx[j]) for j in range(DIM)] for i in range(DIM)]
    ReUdD = [[sp.diff(ReU[i], xx[j]) for j in range(DIM)] for i in range(DIM)]
    ReUdDD = [[[sp.diff(ReUdD[i][j], xx[k]) for k in range(DIM)] for j in range(DIM)] for i in range(DIM)]
    ReDDdD = [[[sp.diff(ReDD[i][j], xx[k]) for k in range(DIM)] for j in range(DIM)] for i in range(DIM)]
    ReDDdDD = [[[[sp.diff(ReDDdD[i][j][k], xx[l]) for l in range(DIM)] for k in range(DIM)]
                for j in range(DIM)] for i in range(DIM)]
    simplify = SymPySimplifyExpressions and not precompute
    ghatDDdD = [[[(sp.simplify(sp.diff(ghatDD[i][j], xx[k])) if simplify else sp.diff(ghatDD[i][j], xx[k]))
                  for k in range(DIM)] for j in range(DIM)] for i in range(DIM)]
    ghatDDdDD = [[[[sp.diff(ghatDDdD[i][j][k], xx[l]) for l in range(DIM)] for k in range(DIM)]
                  for j in range(DIM)] for i in range(DIM)]
    GammahatUDD = ixp.zerorank3()
    for i in range(DIM):
        for k in range(DIM):
            for l in range(DIM):
                for m in range(DIM):
                    GammahatUDD[i][k][l] += sp.Rational(1, 2) * ghatUU[i][m] * (
                        ghatDDdD[m][k][l] + ghatDDdD[m][l][k] - ghatDDdD[k][l][m])
    GammahatUDDdD = [[[[sp.diff(GammahatUDD[i][j][k], xx[l]) for l in range(DIM)] for k in range(DIM)]
                      for j in range(DIM)] for i in range(DIM)]
    if not precompute:
        return

    names = ["ReU", "ReDD", "ghatDD", "ghatUU", "detgammahat", "detgammahatdD", "detgammahatdDD",
             "ReUdD", "ReUdDD", "ReDDdD", "ReDDdDD", "ghatDDdD", "ghatDDdDD", "GammahatUDD", "GammahatUDDdD"]
    g = globals()
    for n in names:
        g[n] = _map_tensor(_to_precompute_symbols, g[n])
    # Find every generic-function symbol, evaluate it in closed form, and fold
    # in those that do not depend on xx (reference_metric.py:692-770).
    free = set()
    _map_tensor(lambda e: free.update(sp.sympify(e).free_symbols), [g[n] for n in names])
    fsyms = sorted((s for s in free if any(str(s).startswith(f) for f in _FUNC_NAMES)), key=str)
    subs = {}
    rfm_precompute_symbols.clear()
    for s in fsyms:
        val = _closed_form_of_precompute_symbol(str(s))
        if not (val.free_symbols & set(xx)):
            subs[s] = val
        else:
            rfm_precompute_symbols[str(s)] = val
    if subs:
        for n in names:
            g[n] = _map_tensor(lambda e: sp.sympify(e).xreplace(subs), g[n])


def rfm_precompute_tables():
    """[(symbol name, closed form in xx, axes it depends on)] for device tables.

    Each table is 1-D along one axis, or 2-D in (xx0,xx1) stored as
    ``i0 + N0*i1`` (reference_metric.py:844-859)."""
    out = []
    for name in sorted(rfm_precompute_symbols):
        expr = rfm_precompute_symbols[name]
        axes = tuple(d for d in range(3) if xx[d] in expr.free_symbols)
        out.append((name, expr, axes))
    return out


def compute_Jacobian_and_inverseJacobian_tofrom_Cartesian():
    """(d xCart^i / d xx^j, its inverse d xx^i / d xCart^j) for the current
    coordinate system (reference_metric.py:904-911)."""
    J = [[sp.diff(xx_to_Cart[i], xx[j]) for j in range(3)] for i in range(3)]
    Jinv, _ = ixp.generic_matrix_inverter3x3(J)
    return J, Jinv


def basis_transform_vectorU_from_rfmbasis_to_Cartesian(Jac_dUCart_dDrfmUD, src_vectorU):
    return [sum(Jac_dUCart_dDrfmUD[i][l] * src_vectorU[l] for l in range(3)) for i in range(3)]


def basis_transform_vectorU_from_Cartesian_to_rfmbasis(Jac_dUrfm_dDCartUD, Cart_src_vectorU):
    return [sum(Jac_dUrfm_dDCartUD[i][l] * Cart_src_vectorU[l] for l in range(3)) for i in range(3)]


def basis_transform_tensorDD_from_rfmbasis_to_Cartesian(Jac_dUrfm_dDCartUD, src_tensorDD):
    return [[sum(Jac_dUrfm_dDCartUD[l][i] * Jac_dUrfm_dDCartUD[m][j] * src_tensorDD[l][m]
                 for l in range(3) for m in range(3)) for j in range(3)] for i in range(3)]


def basis_transform_tensorDD_from_Cartesian_to_rfmbasis(Jac_dUCart_dDrfmUD, Cart_src_tensorDD):
    return [[sum(Jac_dUCart_dDrfmUD[l][i] * Jac_dUCart_dDrfmUD[m][j] * Cart_src_tensorDD[l][m]
                 for l in range(3) for m in range(3)) for j in range(3)] for i in range(3)]


def get_EigenCoord():
    """Coordinate family used to build boundary-condition maps (reference_metric.py:946-952)."""
    cs = par.parval_from_str("reference_metric::CoordSystem")
    for e in ("Spherical", "Cylindrical", "SymTP", "Cartesian"):
        if e in cs:
            return e
    raise ValueError("no EigenCoord for CoordSystem " + cs)


def default_xxmin_xxmax_numeric(CoordSystem, params):
    """Numeric xxmin/xxmax given a parameter dict (the C version of
    set_Nxx_dxx_invdx_params__and__xx, reference_metric.py:1330-1395)."""
    import math
    pi = math.pi
    if CoordSystem == "Spherical":
        return [0.0, 0.0, -pi], [params.get("RMAX", 10.0), pi, pi]
    if CoordSystem in ("SinhSpherical", "SinhSphericalv2"):
        return [0.0, 0.0, -pi], [1.0, pi, pi]
    if CoordSystem == "Cylindrical":
        return [0.0, -pi, params.get("ZMIN", -10.0)], [params.get("RHOMAX", 10.0), pi, params.get("ZMAX", 10.0)]
    if CoordSystem in ("SinhCylindrical", "SinhCylindricalv2"):
        return [0.0, -pi, -1.0], [1.0, pi, 1.0]
    if CoordSystem == "Cartesian":
        return ([params.get("xmin", -10.0), params.get("ymin", -10.0), params.get("zmin", -10.0)],
                [params.get("xmax", 10.0), params.get("ymax", 10.0), params.get("zmax", 10.0)])
    raise ValueError(CoordSystem)
