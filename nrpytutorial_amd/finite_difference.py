"""Finite-difference code generation: ``FD_outputC``.

Same interface and algorithm as finite_difference.py / finite_difference_helpers.py:

1. derivative discovery: every free symbol of the right-hand sides that is
   neither a gridfunction nor a C parameter and contains ``_dD``, ``_dDD``,
   ``_dupD``, ``_ddnD`` or ``_dKOD`` is a derivative; with an upwind control
   vector each ``_dupD`` gains a ``_ddnD`` twin (finite_difference_helpers.py:24-98);
2. the name splits into base gridfunction and operator, e.g.
   ``hDD_dDD0112`` -> (``hDD01``, ``dDD12``) (finite_difference_helpers.py:105-163);
3. exact rational stencil weights from a Vandermonde inverse
   (``compute_fdcoeffs_fdstencl``, finite_difference.py:252-369);
4. every needed point is read once, in memory-address order
   (finite_difference_helpers.py:291-460);
5. stencils are CSE'd with prefix FDPart1, upwinded derivatives become
   ``UpWind_i*(dupD - ddnD) + ddnD`` with ``UpWind_i = (beta^i > 0) ? 1 : 0``
   (FDPart2, finite_difference_helpers.py:729-820; SIMD/SIMD_intrinsics.h:119-124),
   and the right-hand sides are CSE'd with prefix FDPart3 and written out
   (finite_difference_helpers.py:822-853).

With ``finite_difference::FD_functions_enable`` set, step 5 calls one function
per finite-difference operator instead of inlining the stencils, as
``add_FD_func_to_outC_function_dict`` does (finite_difference_helpers.py:508-593):
``order_<N>_f_<op>(invdx..., f_<point>...)``, registered in ``fd_function_dict``
and written out by ``output_finite_difference_functions_h``. Only the upwind
control vector stays inline. The functions are ``NRPY_FD_FUNC``-qualified:
``static __device__ __forceinline__`` in the HIP bundles (a call per stencil
would spill the ~240 live stencil values of a BSSN lane), ``static inline`` in
plain C.

The result is a per-gridpoint block of C statements that is valid HIP device
code; ``loop.simple_loop`` / ``outputC.outCfunction`` place it in a gfx950
kernel where one lane owns one gridpoint.
"""
from collections import OrderedDict
from fractions import Fraction
import os
import sympy as sp

from . import NRPy_param_funcs as par
from . import grid as gri
from .outputC import outputC, parse_outCparams_string, superfast_uniq

modulename = "finite_difference"
par.initialize_param(par.glb_param("int", modulename, "FD_CENTDERIVS_ORDER", 4))
par.initialize_param(par.glb_param("bool", modulename, "FD_functions_enable", "False"))
par.initialize_param(par.glb_param("int", modulename, "FD_KO_ORDER__CENTDERIVS_PLUS", 2))
# Gridfunctions whose centred stencils are summed as differences of nearby values (no
# counterpart in the reference; comma-separated names, default none). The weights of a
# centred first derivative are odd, those of a centred second derivative even and summing
# to zero, those of a mixed second derivative odd in each direction, so
#   dD:  sum_j w_j (f(+j) - f(-j)),
#   dDD: sum_j w_j ((f(+j) - f(0)) + (f(-j) - f(0))),
#   mixed dDD: sum_{a,b>0} w_a w_b ((f(a,b) - f(a,-b)) - (f(-a,b) - f(-a,-b)))
# equal the plain weighted sums. For a field with a large constant part (the BSSN lapse
# and conformal factor, ~1) each inner difference is exact in floating point (Sterbenz:
# the two values within a factor 2) and small, so the O(1) part no longer leaves rounding
# of ~1e-16/dx^2 in second derivatives. Upwinded and dissipation stencils are unchanged.
par.initialize_param(par.glb_param("char", modulename, "FD_difference_from_centre", ""))

DERIV_TAGS = ("_dD", "_dupD", "_ddnD", "_dKOD")

# FD_functions_enable: function name -> C definition (finite_difference_helpers.py:508-593)
fd_function_dict = OrderedDict()

FD_FUNC_QUALIFIER = """#ifndef NRPY_FD_FUNC
#ifdef __HIPCC__
#define NRPY_FD_FUNC static __device__ __forceinline__
#else
#define NRPY_FD_FUNC static inline
#endif
#endif
"""


def _fd_functions_enabled():
    v = par.parval_from_str("finite_difference::FD_functions_enable")
    return v is True or str(v) == "True"


def _fd_function_call(lhs, base, op, coeffs, stencl):
    """Register the operator function for ``op`` (once) and return the statement that calls it
    for the gridfunction ``base`` (finite_difference_helpers.py:508-593): arguments are the
    invdx factors the operator uses, then the stencil points in stencil order."""
    order = par.parval_from_str("finite_difference::FD_CENTDERIVS_ORDER")
    fname = "order_%d_f_%s" % (order, op)
    dirs = [int(op[-2]), int(op[-1])] if op.startswith("dDD") else [int(op[-1])]
    used = sorted(set(dirs))
    if fname not in fd_function_dict:
        invdx = [sp.Symbol("invdx%d" % d, real=True) for d in range(3)]
        e = sp.sympify(0)
        for c, st in zip(coeffs, stencl):
            e += c * sp.Symbol("f" + _suffix(st))
        for d in dirs:
            e *= invdx[d]
        body = outputC(e, "retval", "returnstring",
                       params="includebraces=False,outCverbose=False,CSE_preprocess=True")
        body = body.replace("retval = ", "return ")
        args = ["const double invdx%d" % d for d in used] + ["const double f" + _suffix(st) for st in stencl]
        desc = (op.replace("dDD", "second derivative: ").replace("dD", "first derivative: ")
                .replace("dKOD", "Kreiss-Oliger derivative: ").replace("dupD", "upwinded derivative: ")
                .replace("ddnD", "downwinded derivative: "))
        fd_function_dict[fname] = ("/* Finite-difference operator, %s direction. */\n"
                                   "NRPY_FD_FUNC double %s(%s) {\n%s}\n"
                                   % (desc, fname, ", ".join(args), "".join("  " + l + "\n" for l in body.splitlines())))
    callargs = ["invdx%d" % d for d in used] + [base + _suffix(st) for st in stencl]
    return "%s = %s(%s);" % (lhs, fname, ", ".join(callargs))


def finite_difference_functions_code():
    """Definitions of every registered operator function (empty if none)."""
    if not fd_function_dict:
        return ""
    return FD_FUNC_QUALIFIER + "".join(fd_function_dict.values())


def output_finite_difference_functions_h(path=os.path.join(".")):
    """Write finite_difference_functions.h (finite_difference.py:159-181)."""
    with open(os.path.join(path, "finite_difference_functions.h"), "w") as f:
        f.write("#ifndef __FD_FUNCTIONS_H__\n#define __FD_FUNCTIONS_H__\n#include <math.h>\n"
                + finite_difference_functions_code() + "#endif // #ifndef __FD_FUNCTIONS_H__\n")


def _solve_weights(offsets, m):
    """Weights w_j with sum_j w_j x_j^n = m! delta_{n,m} for n < len(offsets), exact."""
    n = len(offsets)
    A = [[Fraction(x) ** r for x in offsets] for r in range(n)]
    b = [Fraction(0)] * n
    fact = 1
    for k in range(2, m + 1):
        fact *= k
    b[m] = Fraction(fact)
    # Gauss-Jordan elimination on exact rationals
    M = [A[r] + [b[r]] for r in range(n)]
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        pv = M[c][c]
        M[c] = [v / pv for v in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [a - f * bb for a, bb in zip(M[r], M[c])]
    return [M[r][n] for r in range(n)]


def compute_fdcoeffs_fdstencl(derivstring, FDORDER=-1):
    """Exact stencil coefficients and offsets for a derivative operator such as
    ``dD0``, ``dDD12``, ``dupD2``, ``ddnD1``, ``dKOD0``.

    Returns (coefficients as SymPy Rationals, offsets as [d0,d1,d2,0] lists),
    without the 1/dx factors."""
    if FDORDER == -1:
        FDORDER = par.parval_from_str("finite_difference::FD_CENTDERIVS_ORDER")
        if "dKOD" in derivstring:
            FDORDER += par.parval_from_str("finite_difference::FD_KO_ORDER__CENTDERIVS_PLUS")
    size = FDORDER + 1
    shift = 0
    if "dupD" in derivstring:
        shift = 1
    elif "ddnD" in derivstring:
        shift = -1
    elif "dfullupD" in derivstring:
        shift = FDORDER // 2
    elif "dfulldnD" in derivstring:
        shift = -(FDORDER // 2)
    offsets = [j - (size - 1) // 2 + shift for j in range(size)]
    if "DDD" in derivstring:
        raise ValueError("only derivatives up to second order are supported")
    kind, row = "first", 1
    if "dKOD" in derivstring:
        kind, row = "KO", size - 1
    elif "DD" in derivstring:
        kind = "second" if derivstring[-1] == derivstring[-2] else "mixed"
        row = 2 if kind == "second" else 1
    w = _solve_weights(offsets, row)
    coeffs, stencl = [], []
    if kind == "mixed":
        d1, d2 = int(derivstring[-1]), int(derivstring[-2])
        for a, wa in zip(offsets, w):
            for b, wb in zip(offsets, w):
                c = wa * wb
                if c != 0:
                    idx = [0, 0, 0, 0]
                    idx[d1] = a
                    idx[d2] = b
                    coeffs.append(sp.Rational(c.numerator, c.denominator))
                    stencl.append(idx)
        return coeffs, stencl
    d = int(derivstring[-1])
    for o, c in zip(offsets, w):
        if kind == "KO":
            c = c * Fraction((-1) ** ((size + 1) // 2), 2 ** row)
        if c != 0:
            idx = [0, 0, 0, 0]
            idx[d] = o
            coeffs.append(sp.Rational(c.numerator, c.denominator))
            stencl.append(idx)
    return coeffs, stencl


def split_deriv_var(varname):
    """``hDD_dDD0112`` -> (``hDD01``, ``dDD12``)."""
    us = varname.rfind("_d")
    if us < 0:
        raise ValueError(varname + " is not a derivative variable")
    left, right = varname[:us], varname[us + 1:]
    rank = sum(1 for c in left if c in "UD")
    digits = len(right) - len(right.rstrip("0123456789"))
    order = sum(1 for c in right if c == "D")
    if rank + order != digits:
        raise ValueError(varname + ": number of U/D letters does not match trailing digits")
    letters = right[:len(right) - digits]
    idx = right[len(right) - digits:]
    return left + idx[:rank], letters + idx[rank:]


def list_deriv_vars(sympyexpr_list, upwindcontrolvec):
    from . import reference_metric as rfm
    found = []
    for e in sympyexpr_list:
        for v in sp.sympify(e.rhs).free_symbols:
            if gri.variable_type(v) == "other":
                if str(v) in rfm.rfm_precompute_symbols:
                    continue  # read from the precomputed reference-metric tables
                if any(t in str(v) for t in DERIV_TAGS):
                    found.append(v)
                else:
                    raise ValueError("unregistered variable '%s' in expression for %s" % (v, e.lhs))
    found = superfast_uniq(found)
    if upwindcontrolvec != "":
        for v in list(found):
            if "_dupD" in str(v):
                found.append(sp.Symbol(str(v).replace("_dupD", "_ddnD")))
    found = superfast_uniq(found)
    return sorted(found, key=sp.default_sort_key)


def _differenced_stencil(b, op, coeffs, stencl, stmts):
    """The stencil of op on gridfunction b as a weighted sum of differences of nearby values
    (FD_difference_from_centre), appending the difference statements to stmts; None for
    stencils that are left as plain sums (upwinded, dissipation) or lack the symmetry."""
    w = {tuple(s[:3]): c for c, s in zip(coeffs, stencl)}
    invdx = [sp.Symbol("invdx%d" % d, real=True) for d in range(3)]
    if op.startswith("dD") and not op.startswith("dDD") and len(op) == 3:
        d = int(op[-1])
        e = sp.sympify(0)
        for off, c in w.items():
            j = off[d]
            if j > 0:
                neg = tuple(-x for x in off)
                if w.get(neg) != -c:
                    return None
                nm = "%s_%s_d%d" % (b, op, j)
                stmts.append("const double %s = %s%s - %s%s;" % (nm, b, _suffix(off), b, _suffix(neg)))
                e += c * sp.Symbol(nm)
        return e * invdx[d]
    if op.startswith("dDD") and op[-1] == op[-2]:
        d = int(op[-1])
        if sum(w.values()) != 0:
            return None
        e = sp.sympify(0)
        for off, c in w.items():
            j = off[d]
            if j > 0:
                neg = tuple(-x for x in off)
                if w.get(neg) != c:
                    return None
                nm = "%s_%s_p%d" % (b, op, j)
                stmts.append("const double %s = (%s%s - %s) + (%s%s - %s);"
                             % (nm, b, _suffix(off), b, b, _suffix(neg), b))
                e += c * sp.Symbol(nm)
        return e * invdx[d] * invdx[d]
    if op.startswith("dDD"):
        d1, d2 = int(op[-2]), int(op[-1])
        e = sp.sympify(0)
        for off, c in w.items():
            a, bb = off[d1], off[d2]
            if a > 0 and bb > 0:
                def pt(sa, sb):
                    o = [0, 0, 0]
                    o[d1], o[d2] = sa, sb
                    return tuple(o)
                if not (w.get(pt(a, -bb)) == -c and w.get(pt(-a, bb)) == -c and w.get(pt(-a, -bb)) == c):
                    return None
                nm = "%s_%s_q%d%d" % (b, op, a, bb)
                stmts.append("const double %s = (%s%s - %s%s) - (%s%s - %s%s);"
                             % (nm, b, _suffix(pt(a, bb)), b, _suffix(pt(a, -bb)),
                                b, _suffix(pt(-a, bb)), b, _suffix(pt(-a, -bb))))
                e += c * sp.Symbol(nm)
        return e * invdx[d1] * invdx[d2]
    return None


def _suffix(idx):
    if list(idx[:3]) == [0, 0, 0]:
        return ""
    parts = []
    for d in range(3):
        o = idx[d]
        parts.append("i%d" % d + ("" if o == 0 else ("p%d" % o if o > 0 else "m%d" % -o)))
    return "_" + "_".join(parts)


def _ijk(idx):
    parts = []
    for d in range(3):
        o = idx[d]
        parts.append("i%d" % d + ("" if o == 0 else ("+%d" % o if o > 0 else "-%d" % -o)))
    return ",".join(parts)


def _read_statement(gfname, idx, array):
    """``const double hDD00_i0m1_i1_i2 = <access>;`` with the access spelled by gri.gfaccess
    (finite_difference_helpers.py:234-238), so grid::GridFuncMemAccess applies to reads too."""
    return "const double %s%s = %s;" % (gfname, _suffix(idx), gri.gfaccess(array, gfname, _ijk(idx)))


def FD_outputC_readblock(sympyexpr_list, upwindcontrolvec=""):
    """Steps 1-2 of FD_outputC for ``sympyexpr_list``: memory reads, finite-difference
    stencils (CSE prefix FDPart1) and upwinding (FDPart2). Returns (C code, number of
    steps emitted, derivative symbols). The adjoint generator reuses it to rebuild the
    forward quantities of a right-hand side before its reverse sweep."""
    if not isinstance(sympyexpr_list, list):
        sympyexpr_list = [sympyexpr_list]
    derivs = list_deriv_vars(sympyexpr_list, upwindcontrolvec)
    bases, ops = [], []
    for v in derivs:
        b, o = split_deriv_var(str(v))
        if gri.variable_type(b) != "gridfunction":
            raise ValueError("derivative of %s, which is not a registered gridfunction" % b)
        if not any(t in o for t in ("dD", "dupD", "ddnD", "dKOD")):
            raise ValueError("no valid derivative operator in " + o)
        bases.append(b)
        ops.append(o)
    stencils = [compute_fdcoeffs_fdstencl(o) for o in ops]
    diff_gfs = set(n.strip() for n in str(par.parval_from_str("finite_difference::FD_difference_from_centre"))
                   .split(",") if n.strip())
    if diff_gfs and _fd_functions_enabled():
        raise ValueError("FD_difference_from_centre applies to inline stencils only; disable FD_functions_enable")

    # Points to read, per gridfunction, ordered by memory address (i2 slowest).
    points = {}
    for b, (c, st) in zip(bases, stencils):
        points.setdefault(b, []).extend(tuple(s) for s in st)
    for e in sympyexpr_list:
        for v in sp.sympify(e.rhs).free_symbols:
            if gri.variable_type(v) == "gridfunction":
                points.setdefault(str(v), []).append((0, 0, 0, 0))
    diffs, diff_expr = [], {}
    for v, b, o, (coeffs, st) in zip(derivs, bases, ops, stencils):
        if b in diff_gfs:
            e = _differenced_stencil(b, o, coeffs, st, diffs)
            if e is not None:
                diff_expr[str(v)] = e
                if o.startswith("dDD") and o[-1] == o[-2]:
                    points[b].append((0, 0, 0, 0))  # the centre value of the pair sums
    order = {n: k for k, n in enumerate(gri.glb_gridfcs_list and [g.name for g in gri.glb_gridfcs_list])}
    reads = []
    for gf in sorted(points, key=lambda n: order.get(n, 0)):
        pts = sorted(set(points[gf]), key=lambda s: (s[2], s[1], s[0]))
        for s in pts:
            reads.append(_read_statement(gf, s, "in_gfs"))
    reads += diffs

    invdx = [sp.Symbol("invdx%d" % d, real=True) for d in range(3)]
    FDexprs, FDnames, funccalls = [], [], []
    use_functions = _fd_functions_enabled()
    for v, b, o, (coeffs, st) in zip(derivs, bases, ops, stencils):
        name = str(v)
        if upwindcontrolvec != "" and ("_dupD" in name or "_ddnD" in name):
            name = "UpwindAlgInput" + name
        if use_functions:
            funccalls.append(_fd_function_call("const double " + name, b, o, coeffs, st))
            continue
        if str(v) in diff_expr:
            FDexprs.append(diff_expr[str(v)])
            FDnames.append("const double " + name)
            continue
        e = sp.sympify(0)
        for c, s in zip(coeffs, st):
            e += c * sp.Symbol(b + _suffix(s))
        if o.startswith("dDD"):
            e *= invdx[int(o[-2])] * invdx[int(o[-1])]
        else:
            e *= invdx[int(o[-1])]
        FDexprs.append(e)
        FDnames.append("const double " + name)

    upwind_dirs = []
    if upwindcontrolvec != "":
        upwind_dirs = sorted(set(int(o[-1]) for o in ops if "dupD" in o))
        for d in upwind_dirs:
            FDexprs.append(upwindcontrolvec[d])
            FDnames.append("const double UpwindControlVectorU%d" % d)

    out = ""
    nsteps = 1 + (1 if reads else 0) + (1 if upwind_dirs else 0)
    step = 1
    if reads:
        out += "/* Step %d of %d: read gridfunctions and compute finite-difference stencils */\n" % (step, nsteps)
        step += 1
        if use_functions:
            out += "\n".join(reads + funccalls) + "\n"
            if FDexprs:
                out += outputC(FDexprs, FDnames, "returnstring",
                               params="includebraces=False,CSE_varprefix=FDPart1,CSE_preprocess=True,"
                                      "outCverbose=False")
        elif FDexprs:
            out += outputC(FDexprs, FDnames, "returnstring",
                           params="includebraces=False,CSE_varprefix=FDPart1,CSE_preprocess=True,outCverbose=False",
                           prestring="\n".join(reads))
        else:
            out += "\n".join(reads) + "\n"
    if upwindcontrolvec != "":
        if upwind_dirs:
            out += "/* Step %d of %d: upwinding */\n" % (step, nsteps)
            step += 1
            for d in upwind_dirs:
                out += "const double UpWind%d = (UpwindControlVectorU%d > 0.0) ? 1.0 : 0.0;\n" % (d, d)
        up_exprs, up_names = [], []
        for v, o in zip(derivs, ops):
            if "dupD" in o:
                up = sp.Symbol("UpwindAlgInput" + str(v))
                dn = sp.Symbol("UpwindAlgInput" + str(v).replace("_dupD", "_ddnD"))
                up_exprs.append(sp.Symbol("UpWind%s" % o[-1]) * (up - dn) + dn)
                up_names.append("const double " + str(v))
        if up_exprs:
            out += outputC(up_exprs, up_names, "returnstring",
                           params="includebraces=False,CSE_varprefix=FDPart2,outCverbose=False")
    return out, step, nsteps, derivs


def FD_outputC(filename, sympyexpr_list, params="", upwindcontrolvec=""):
    """Generate the per-gridpoint block for a list of ``lhrh`` pairs."""
    p = parse_outCparams_string(params)
    if not isinstance(sympyexpr_list, list):
        sympyexpr_list = [sympyexpr_list]
    indent = "  " if p.includebraces == "True" else ""
    fullindent = indent + p.preindent
    out, step, nsteps, _ = FD_outputC_readblock(sympyexpr_list, upwindcontrolvec)
    out += "/* Step %d of %d: evaluate right-hand sides and write to memory */\n" % (step, nsteps)
    out += outputC([e.rhs for e in sympyexpr_list], [e.lhs for e in sympyexpr_list], "returnstring",
                   params="includebraces=False,CSE_varprefix=FDPart3,CSE_preprocess=True,outCverbose=False")
    out = "".join(fullindent + l + "\n" for l in out.splitlines())
    if p.includebraces == "True":
        out = p.preindent + "{\n" + out + p.preindent + "}\n"
    if filename == "stdout":
        print(out)
    elif filename == "returnstring":
        return out
    else:
        with open(filename, p.outCfileaccess) as f:
            f.write(out)
    return None
