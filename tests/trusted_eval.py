"""Evaluate symbolic globals the way the reference's unit tests do, and compare
them with the reference's trusted values.

Scheme (UnitTesting/cse_simplify_and_evaluate_sympy_expressions.py:56-101,
UnitTesting/calc_error.py:59-67, UnitTesting/standard_constants.py):

* every free symbol ``s`` gets ``mpf(random.random())`` after
  ``random.seed(int(md5(str(s)).hexdigest(), 16))``; ``M_PI`` is pi and
  ``M_SQRT1_2`` is 1/sqrt(2);
* expressions are evaluated at 30 significant digits; a result smaller than
  1e-20 is re-evaluated at 60 digits and set to exactly 0 if it stays below
  1e-40;
* a value passes when log10 of its relative error is below -15 (absolute
  error when the trusted value is 0).
"""
import hashlib
import json
import os
import random

import mpmath
import sympy as sp

PRECISION = 30
HERE = os.path.dirname(os.path.abspath(__file__))


def load_trusted(entry):
    with open(os.path.join(HERE, "golden", "trusted_values.json")) as f:
        d = json.load(f)
    return d[entry]["values"]


def flatten_globals(name_to_obj):
    """{'hDD': [[..]..]} -> {'hDD[0][0]': expr, ...} (nested lists indexed like the reference)."""
    out = {}

    def walk(prefix, obj):
        if isinstance(obj, (list, tuple)):
            for i, o in enumerate(obj):
                walk(prefix + "[%d]" % i, o)
        else:
            out[prefix] = sp.sympify(obj)
    for k, v in name_to_obj.items():
        walk(k, v)
    return out


def seeded_value(sym):
    name = str(sym)
    if name == "M_PI":
        return mpmath.pi
    if name == "M_SQRT1_2":
        return 1 / mpmath.sqrt(2)
    random.seed(int(hashlib.md5(name.encode()).hexdigest(), 16))
    return mpmath.mpf(random.random())


def _evaluate(expr, dps):
    with mpmath.workdps(dps):
        free = sorted(expr.free_symbols, key=str)
        vals = {s: seeded_value(s) for s in free}
        repl, red = sp.cse([expr], order="none")
        env = dict(vals)
        nrpyAbs = sp.Function("nrpyAbs")
        for lhs, rhs in repl:
            rhs = rhs.replace(nrpyAbs, sp.Abs)
            f = sp.lambdify(list(rhs.free_symbols), rhs, modules="mpmath")
            env[lhs] = f(*[env[s] for s in rhs.free_symbols])
        final = red[0].replace(nrpyAbs, sp.Abs)
        fs = list(final.free_symbols)
        f = sp.lambdify(fs, final, modules="mpmath")
        return f(*[env[s] for s in fs])


def evaluate(expr):
    expr = sp.sympify(expr)
    if expr.free_symbols == set():
        with mpmath.workdps(PRECISION):
            v = complex(expr.evalf(PRECISION + 10)) if not expr.is_real else None
            return mpmath.mpf(str(expr.evalf(PRECISION + 10))) if expr.is_real is not False else mpmath.mpc(v)
    val = _evaluate(expr, PRECISION)
    with mpmath.workdps(PRECISION):
        if val != 0 and abs(val) < mpmath.mpf(10) ** (-2.0 / 3 * PRECISION):
            val2 = _evaluate(expr, 2 * PRECISION)
            if abs(val2) < mpmath.mpf(10) ** (-4.0 / 3 * PRECISION):
                return mpmath.mpf(0)
    return val


def _to_mp(v):
    if isinstance(v, list):
        return mpmath.mpc(mpmath.mpf(v[0]), mpmath.mpf(v[1]))
    return mpmath.mpf(v)


def compare(entry, name_to_obj, ignore_missing=()):
    """Return a list of failure strings (empty when everything agrees)."""
    trusted = load_trusted(entry)
    computed = flatten_globals(name_to_obj)
    failures = []
    for key in sorted(trusted):
        if key not in computed:
            if not any(key.startswith(p) for p in ignore_missing):
                failures.append("missing " + key)
            continue
        with mpmath.workdps(PRECISION):
            t = _to_mp(trusted[key])
            c = evaluate(computed[key])
            if t == 0:
                err = abs(c)
            elif c == 0:
                err = abs(t)
            else:
                err = abs((t - c) / t)
            if err != 0 and mpmath.log10(err) >= -PRECISION / 2.0:
                failures.append("%s: trusted %s computed %s" % (key, mpmath.nstr(t, 20), mpmath.nstr(c, 20)))
    return failures
