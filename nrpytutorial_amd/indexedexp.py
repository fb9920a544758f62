"""Indexed expressions (tensors as nested lists of SymPy expressions).

Interface of indexedexp.py (indexedexp.py:13-508):

* ``declarerank1..4(symbol, symmetry)`` create tensors of symbols named
  ``symbol + digits``; a symmetry such as ``"sym01"`` or ``"sym01_sym23"``
  makes the listed index pairs interchangeable, so ``hDD[1][0]`` is the
  symbol ``hDD01`` (indexedexp.py:84-168);
* the codegen-time parameter ``indexedexp::symmetry_axes`` names axes along
  which the solution is constant; every declared derivative symbol that
  differentiates along such an axis is replaced by 0 (indexedexp.py:189-251);
* closed-form inverses of symmetric and generic 3x3 (and 4x4 symmetric)
  matrices (indexedexp.py:331-470).
"""
import itertools
import sympy as sp
from . import NRPy_param_funcs as par
from . import grid as gri

thismodule = "indexedexp"
par.initialize_param(par.glb_param("char", thismodule, "symmetry_axes", ""))


def _parse_symmetry(symmetry, rank):
    """List of (i, j, sign) index-pair relations from a symmetry string."""
    rels = []
    if not symmetry or symmetry == "nosym":
        return rels
    for part in symmetry.split("_"):
        if part == "nosym":
            continue
        if part.startswith("sym"):
            sign, idx = 1, part[3:]
        elif part.startswith("anti"):
            sign, idx = -1, part[4:]
        else:
            raise ValueError("unsupported symmetry option '" + part + "'")
        digits = [int(c) for c in idx]
        if len(digits) < 2 or max(digits) >= rank:
            raise ValueError("bad symmetry '" + part + "' for rank %d" % rank)
        for a, b in itertools.combinations(digits, 2):
            rels.append((a, b, sign))
    return rels


def _canonical(index, rels):
    """Canonical index tuple and sign, by repeatedly ordering related pairs."""
    index = list(index)
    sign = 1
    changed = True
    while changed:
        changed = False
        for a, b, s in rels:
            if index[b] < index[a]:
                index[a], index[b] = index[b], index[a]
                sign *= s
                changed = True
            elif index[a] == index[b] and s < 0:
                return None, 0
    return tuple(index), sign


def _derivative_crosses_symmetry_axis(name, axes):
    """True if symbol ``name`` is a derivative taken along one of ``axes``."""
    pos = name.find("_d")
    if pos < 0:
        return False
    order = name[pos + 2:].count("D")
    if order > 2:
        raise ValueError("derivative order > 2 not supported: " + name)
    tail = name[len(name) - order:] if order > 0 else ""
    return any(c in axes for c in tail)


def _zero_symmetry_axis_derivs(obj):
    axes = par.parval_from_str("indexedexp::symmetry_axes")
    if axes == "":
        return obj

    def walk(o):
        if isinstance(o, list):
            return [walk(x) for x in o]
        if isinstance(o, sp.Symbol) and _derivative_crosses_symmetry_axis(str(o), axes):
            return sp.sympify(0)
        if isinstance(o, sp.Mul):
            # signed symbol from an antisymmetric declaration
            syms = [a for a in o.args if isinstance(a, sp.Symbol)]
            if syms and _derivative_crosses_symmetry_axis(str(syms[0]), axes):
                return sp.sympify(0)
        return o
    return walk(obj)


def declare_indexedexp(rank, symbol=None, symmetry=None, dimension=None):
    if dimension is None or dimension == -1:
        dimension = par.parval_from_str("grid::DIM")
    if not isinstance(dimension, int) or dimension <= 0:
        raise ValueError("dimension must be a positive integer")
    rels = _parse_symmetry(symmetry, rank) if symbol is not None else []

    def build(prefix):
        if len(prefix) == rank:
            if symbol is None:
                return sp.sympify(0)
            canon, sign = _canonical(prefix, rels)
            if canon is None:
                return sp.sympify(0)
            s = sp.Symbol(symbol + "".join(str(i) for i in canon))
            return s if sign > 0 else -s
        return [build(prefix + [i]) for i in range(dimension)]
    obj = build([])
    return _zero_symmetry_axis_derivs(obj)


def zerorank1(DIM=-1):
    return declare_indexedexp(1, dimension=DIM)


def zerorank2(DIM=-1):
    return declare_indexedexp(2, dimension=DIM)


def zerorank3(DIM=-1):
    return declare_indexedexp(3, dimension=DIM)


def zerorank4(DIM=-1):
    return declare_indexedexp(4, dimension=DIM)


def declarerank1(symbol, DIM=-1):
    return declare_indexedexp(1, symbol, None, DIM)


def declarerank2(symbol, symmetry, DIM=-1):
    return declare_indexedexp(2, symbol, symmetry, DIM)


def declarerank3(symbol, symmetry, DIM=-1):
    return declare_indexedexp(3, symbol, symmetry, DIM)


def declarerank4(symbol, symmetry, DIM=-1):
    return declare_indexedexp(4, symbol, symmetry, DIM)


def register_gridfunctions_for_single_rank1(gf_type, gf_basename, DIM=-1):
    gri.verify_gridfunction_basename_is_valid(gf_basename)
    obj = declarerank1(gf_basename, DIM)
    gri.register_gridfunctions(gf_type, [str(s) for s in obj], rank=1, is_indexed=True,
                               DIM=len(obj))
    return obj


def register_gridfunctions_for_single_rank2(gf_type, gf_basename, symmetry_option, DIM=-1):
    gri.verify_gridfunction_basename_is_valid(gf_basename)
    obj = declarerank2(gf_basename, symmetry_option, DIM)
    names = []
    for row in obj:
        for s in row:
            if str(s) not in names:
                names.append(str(s))
    gri.register_gridfunctions(gf_type, names, rank=2, is_indexed=True, DIM=len(obj))
    return obj


class NonInvertibleMatrixError(ZeroDivisionError):
    """Matrix not invertible."""


def symm_matrix_inverter3x3(a):
    """Inverse and determinant of a symmetric 3x3 matrix via cofactors."""
    c00 = a[1][1] * a[2][2] - a[1][2] ** 2
    c01 = a[0][2] * a[1][2] - a[0][1] * a[2][2]
    c02 = a[0][1] * a[1][2] - a[0][2] * a[1][1]
    c11 = a[0][0] * a[2][2] - a[0][2] ** 2
    c12 = a[0][1] * a[0][2] - a[0][0] * a[1][2]
    c22 = a[0][0] * a[1][1] - a[0][1] ** 2
    # Determinant as the fully expanded Leibniz sum (six terms, two pairs equal by
    # symmetry). Kept expanded rather than as the cofactor row expansion: for
    # symbolic entries with fractional powers, (det)^(-1/6)-type expressions take
    # their branch from the exact expression tree, and the expanded polynomial is
    # the form the reference's trusted values were computed from
    # (indexedexp.py:331-336).
    det = (a[0][0] * a[1][1] * a[2][2] + 2 * a[0][1] * a[0][2] * a[1][2]
           - a[0][0] * a[1][2] ** 2 - a[1][1] * a[0][2] ** 2 - a[2][2] * a[0][1] ** 2)
    if det == 0:
        raise NonInvertibleMatrixError("matrix has determinant zero")
    inv = [[c00 / det, c01 / det, c02 / det],
           [c01 / det, c11 / det, c12 / det],
           [c02 / det, c12 / det, c22 / det]]
    return inv, det


def symm_matrix_inverter2x2(a):
    det = a[0][0] * a[1][1] - a[0][1] ** 2
    if det == 0:
        raise NonInvertibleMatrixError("matrix has determinant zero")
    return [[a[1][1] / det, -a[0][1] / det], [-a[0][1] / det, a[0][0] / det]], det


def generic_matrix_inverter3x3(a):
    """Inverse and determinant of a general 3x3 matrix via the adjugate."""
    M = sp.Matrix(3, 3, lambda i, j: a[i][j])
    det = (a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
           - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
           + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]))
    if det == 0:
        raise NonInvertibleMatrixError("matrix has determinant zero")
    adj = M.adjugate()
    return [[adj[i, j] / det for j in range(3)] for i in range(3)], det


def symm_matrix_inverter4x4(a):
    M = sp.Matrix(4, 4, lambda i, j: a[i][j])
    det = M.det(method="berkowitz")
    if det == 0:
        raise NonInvertibleMatrixError("matrix has determinant zero")
    adj = M.adjugate()
    return [[adj[i, j] / det for j in range(4)] for i in range(4)], det


def LeviCivitaSymbol_dim3_rank3():
    out = zerorank3(DIM=3)
    for i, j, k in itertools.product(range(3), repeat=3):
        out[i][j][k] = sp.Rational((i - j) * (j - k) * (k - i), 2)
    return out
