"""Expression output: CSE + C/HIP statement printing, and C-function assembly.

Interface of outputC.py:

* ``outputC(sympyexpr, output_varname_str, filename, params, prestring, poststring)``
  common-subexpression-eliminates a list of expressions and prints
  ``const double tmp = ...;`` / ``out = ...;`` statements (outputC.py:208-423);
* ``Pow`` is lowered as the reference does (outputC.py:32-49): x^2..x^5 become
  products, x^-1..x^-5 reciprocals of products, x^(1/2) ``sqrt``,
  x^(1/3) ``cbrt``; ``nrpyAbs`` prints as ``fabs``;
* rational constants are hoisted into named ``const double`` values before CSE
  (the ``CSE_preprocess`` step of cse_helpers.py:16-159);
* ``outCfunction`` / ``add_to_Cfunction_dict`` wrap a per-point body into a
  function (outputC.py:430-487). Here that function is a gfx950 ``__global__``
  kernel plus an ``extern "C"`` launcher; see ``loop.simple_loop``.

The printed statements are the common subset of C and HIP device code, so the
same text compiles for the GPU kernels.
"""
from collections import namedtuple, OrderedDict
import re
import sympy as sp
from sympy.printing.c import C99CodePrinter

from . import NRPy_param_funcs as par

lhrh = namedtuple("lhrh", "lhs rhs")
nrpyAbs = sp.Function("nrpyAbs")

outCparams = namedtuple("outCparams", "preindent includebraces declareoutputvars outCfileaccess outCverbose "
                        "CSE_enable CSE_varprefix CSE_sorting CSE_preprocess SIMD_enable SIMD_find_more_subs "
                        "SIMD_find_more_FMAsFMSs SIMD_debug enable_TYPE gridsuffix")

# Registry of generated functions, name -> dict(kernel=..., launcher=..., meta=...)
outC_function_dict = OrderedDict()
outC_function_prototype_dict = OrderedDict()


def parse_outCparams_string(params):
    """Parse an ``outputC`` option string such as ``"outCverbose=False,CSE_varprefix=FDPart1"``
    (outputC.py:102-202). SIMD options are accepted and ignored: every lane of a
    gfx950 wavefront already works on its own gridpoint."""
    vals = dict(preindent="", includebraces="True", declareoutputvars="False", outCfileaccess="w",
                outCverbose="True", CSE_enable="True", CSE_varprefix="tmp", CSE_sorting="canonical",
                CSE_preprocess="False", SIMD_enable="False", SIMD_find_more_subs="False",
                SIMD_find_more_FMAsFMSs="True", SIMD_debug="False", enable_TYPE="True", gridsuffix="")
    if params:
        for item in params.replace(" ", "").split(","):
            if not item:
                continue
            if "=" not in item:
                raise ValueError("bad outputC option '" + item + "'")
            k, v = item.split("=", 1)
            if k not in vals:
                raise ValueError("unknown outputC option '" + k + "'")
            if k == "preindent":
                v = "   " * int(v)
            vals[k] = v
    return outCparams(**vals)


class _Printer(C99CodePrinter):
    """C99 printer with the reference's Pow and Abs conventions, double literals only."""

    def __init__(self):
        super().__init__({"user_functions": {"nrpyAbs": "fabs"}})

    def _print_Pow(self, expr):
        b, e = expr.as_base_exp()
        bs = self._print(b)
        if e == sp.S.Half:
            return "sqrt(%s)" % bs
        if e == -sp.S.Half:
            return "(1.0/sqrt(%s))" % bs
        if e == sp.Rational(1, 3):
            return "cbrt(%s)" % bs
        if e == -sp.Rational(1, 3):
            return "(1.0/cbrt(%s))" % bs
        if e.is_Integer and 2 <= int(e) <= 5:
            return "(" + "*".join("(%s)" % bs for _ in range(int(e))) + ")"
        if e.is_Integer and -5 <= int(e) <= -1:
            return "(1.0/(" + "*".join("(%s)" % bs for _ in range(-int(e))) + "))"
        return "pow(%s, %s)" % (bs, self._print(e))

    def _print_Rational(self, expr):
        return "(%d.0/%d.0)" % (expr.p, expr.q)

    def _print_Integer(self, expr):
        return "%d" % int(expr)

    def _print_Float(self, expr):
        return repr(float(expr))

    def _print_Symbol(self, expr):
        return str(expr.name)


_printer = _Printer()


def ccode(expr):
    return _printer.doprint(expr)


def _rational_name(prefix, r):
    sgn = "m" if r < 0 else ""
    return "%s_Rational_%s%d_%d" % (prefix, sgn, abs(r.p), r.q)


def cse_preprocess(exprs, prefix=""):
    """Replace non-integer rationals by named symbols so that CSE treats them as
    leaves; returns (exprs, map). The constants print as exact ``p.0/q.0`` double
    divisions. A rational that is the exponent of a power stays in place
    (cse_helpers.py:67-68), so x^(1/2) and x^(1/3) still print as sqrt and cbrt
    (``pow`` with a fractional exponent is NaN for a negative base, cbrt is not)."""
    symmap = {}

    def sym(r):
        s = sp.Symbol(_rational_name(prefix or "tmp", r))
        symmap[s] = r
        return s

    def rebuild(e):
        if isinstance(e, sp.Rational):
            return e if isinstance(e, sp.Integer) else sym(e)
        if isinstance(e, sp.Pow):
            ex = e.exp if isinstance(e.exp, sp.Rational) else rebuild(e.exp)
            return sp.Pow(rebuild(e.base), ex, evaluate=False)
        if not e.args or isinstance(e, (sp.Symbol, sp.NumberSymbol)):
            return e
        return e.func(*[rebuild(a) for a in e.args])
    return [rebuild(sp.sympify(e)) for e in exprs], symmap


def outputC(sympyexpr, output_varname_str, filename="stdout", params="", prestring="", poststring=""):
    """Print statements assigning CSE'd expressions to ``output_varname_str``."""
    p = parse_outCparams_string(params)
    single = not isinstance(sympyexpr, list)
    exprs = [sympyexpr] if single else list(sympyexpr)
    names = [output_varname_str] if single else list(output_varname_str)
    if len(exprs) != len(names):
        raise ValueError("outputC: %d expressions but %d output names" % (len(exprs), len(names)))
    exprs = [sp.sympify(e) for e in exprs]
    ind = p.preindent
    TYPE = par.parval_from_str("PRECISION")
    out = ""
    if p.includebraces == "True":
        out += ind + "{\n"
        ind2 = ind + "  "
    else:
        ind2 = ind
    out += "".join(ind2 + l + "\n" for l in prestring.splitlines()) if prestring else ""
    outtype = ""
    if p.declareoutputvars == "True":
        outtype = "const " + TYPE + " "
    if p.CSE_enable == "False":
        for e, n in zip(exprs, names):
            out += ind2 + outtype + n + " = " + ccode(e) + ";\n"
    else:
        prefix = "" if p.CSE_varprefix == "tmp" else p.CSE_varprefix
        if p.CSE_preprocess == "True":
            exprs, ratmap = cse_preprocess(exprs, prefix)
            for s in sorted(ratmap, key=str):
                r = ratmap[s]
                out += ind2 + "const %s %s = %s;\n" % (TYPE, s, "%d.0/%d.0" % (r.p, r.q))
        repl, reduced = sp.cse(exprs, sp.numbered_symbols(p.CSE_varprefix + "_"), order=p.CSE_sorting)
        for lhs, rhs in repl:
            out += ind2 + "const %s %s = %s;\n" % (TYPE, lhs, ccode(rhs))
        for e, n in zip(reduced, names):
            out += ind2 + outtype + n + " = " + ccode(e) + ";\n"
    out += "".join(ind2 + l + "\n" for l in poststring.splitlines()) if poststring else ""
    if p.includebraces == "True":
        out += ind + "}\n"
    if filename == "stdout":
        print(out)
    elif filename == "returnstring":
        return out
    else:
        with open(filename, p.outCfileaccess) as f:
            f.write(out)
    return None


def superfast_uniq(seq):
    seen = set()
    return [x for x in seq if not (x in seen or seen.add(x))]


def check_if_string__error_if_not(allegedstring, stringdesc):
    if not isinstance(allegedstring, str):
        raise TypeError(str(stringdesc) + " == " + str(allegedstring) + " is not a string")


# ---------------------------------------------------------------------------
# C-function assembly -> HIP kernels
# ---------------------------------------------------------------------------

def _parse_params(params):
    """Split a reference-style C parameter list into (ctype, name) pairs."""
    out = []
    for item in params.replace("\n", " ").split(","):
        item = item.strip()
        if not item:
            continue
        m = re.match(r"(.*?)([A-Za-z_][A-Za-z_0-9]*)\s*(\[[0-9]*\])?$", item)
        if not m:
            raise ValueError("cannot parse parameter '" + item + "'")
        ctype = m.group(1).replace("restrict", "").strip()
        name = m.group(2)
        if m.group(3):
            ctype += " " + m.group(3)
        out.append((re.sub(r"\s+", " ", ctype), name))
    return out


def Cfunction(includes=None, prefunc="", desc="", c_type="void", name="", params="", preloop="", body="",
              loopopts="", postloop="", opts="", rel_path_to_Cparams="./", enableCparameters=True):
    """Return (prototype, HIP source) for a per-point function (outputC.py:430-468)."""
    from . import loop as lp
    args = _parse_params(params)
    kernel, launcher, proto = lp.kernel_and_launcher(name=name, desc=desc, args=args, preloop=preloop,
                                                    body=body, loopopts=loopopts, postloop=postloop)
    src = prefunc + "\n" + kernel + "\n" + launcher
    return proto, src


def add_to_Cfunction_dict(includes=None, prefunc="", desc="", c_type="void", name="", params="", preloop="",
                          body="", loopopts="", postloop="", opts="", rel_path_to_Cparams="./", **kwargs):
    proto, src = Cfunction(includes, prefunc, desc, c_type, name, params, preloop, body, loopopts, postloop,
                           opts, rel_path_to_Cparams)
    outC_function_dict[name] = src
    outC_function_prototype_dict[name] = proto


def outCfunction(outfile="", includes=None, prefunc="", desc="", c_type="void", name="", params="", preloop="",
                 body="", loopopts="", postloop="", opts="", rel_path_to_Cparams="./", **kwargs):
    """Write the generated kernel + launcher for ``name`` to ``outfile`` and
    register it (outputC.py:478-487)."""
    if "type" in kwargs:
        c_type = kwargs.pop("type")
    add_to_Cfunction_dict(includes, prefunc, desc, c_type, name, params, preloop, body, loopopts, postloop,
                          opts, rel_path_to_Cparams)
    if outfile and outfile != "returnstring":
        with open(outfile, "w") as f:
            f.write(outC_function_dict[name])
    return outC_function_dict[name]
