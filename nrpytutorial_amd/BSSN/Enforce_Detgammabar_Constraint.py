"""det(gammabar) = det(ghat) enforcement (BSSN/Enforce_Detgammabar_Constraint.py:16-75).

h'_ij = (|det ghat| / det gammabar)^{1/3} (delta_ij + h_ij) - delta_ij, applied at
all points after each RK substep.
"""
import sympy as sp
from .. import grid as gri
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from ..outputC import nrpyAbs, lhrh
from . import BSSN_quantities as Bq


def Enforce_Detgammabar_Constraint_symb_expressions():
    rfm.reference_metric()
    Bq.declare_BSSN_gridfunctions_if_not_declared_already()
    h = Bq.hDD
    Bq.BSSN_basic_tensors()
    _, detgammabar = ixp.symm_matrix_inverter3x3(Bq.gammabarDD)
    out = []
    for i, j in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)):
        delta = 1 if i == j else 0
        hprime = (nrpyAbs(rfm.detgammahat) / detgammabar) ** sp.Rational(1, 3) * (delta + h[i][j]) - delta
        out.append(lhrh(lhs=gri.gfaccess("in_gfs", "hDD%d%d" % (i, j)), rhs=hprime))
    return out
