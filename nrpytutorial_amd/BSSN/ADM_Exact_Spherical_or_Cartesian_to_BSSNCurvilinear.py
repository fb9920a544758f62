"""Exact ADM data given in Spherical or Cartesian coordinates -> rescaled BSSN
variables in the reference-metric basis
(BSSN/ADM_Exact_Spherical_or_Cartesian_to_BSSNCurvilinear.py:16-135).

Jacobian J^a_j = d x^a_{Sph/Cart} / d xx^j; gamma_ij = J^a_i J^b_j gamma_ab,
K_ij likewise, beta^i = (J^{-1})^i_a beta^a.
"""
import itertools
import sympy as sp
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from . import BSSN_in_terms_of_ADM as BitoA

R3 = range(3)


def Convert_Spherical_or_Cartesian_ADM_to_BSSN_curvilinear(CoordType_in, Sph_r_th_ph_or_Cart_xyz,
                                                          gammaDD_inSphorCart, KDD_inSphorCart,
                                                          alpha_inSphorCart, betaU_inSphorCart, BU_inSphorCart):
    if not rfm.have_already_called_reference_metric_function:
        raise RuntimeError("call reference_metric() first")
    if CoordType_in == "Spherical":
        target = rfm.xxSph
    elif CoordType_in == "Cartesian":
        target = rfm.xx_to_Cart
    else:
        raise ValueError("only Spherical or Cartesian ADM input is supported")

    def to_xx(e):
        e = sp.sympify(e)
        return e.subs({Sph_r_th_ph_or_Cart_xyz[i]: target[i] for i in R3}, simultaneous=True)

    alpha = to_xx(alpha_inSphorCart)
    betaS = [to_xx(betaU_inSphorCart[i]) for i in R3]
    BS = [to_xx(BU_inSphorCart[i]) for i in R3]
    gS = [[to_xx(gammaDD_inSphorCart[i][j]) for j in R3] for i in R3]
    KS = [[to_xx(KDD_inSphorCart[i][j]) for j in R3] for i in R3]
    Jac = [[sp.diff(target[i], rfm.xx[j]) for j in R3] for i in R3]
    Jinv, _ = ixp.generic_matrix_inverter3x3(Jac)
    betaU = [sum(Jinv[i][j] * betaS[j] for j in R3) for i in R3]
    BU = [sum(Jinv[i][j] * BS[j] for j in R3) for i in R3]
    gammaDD = ixp.zerorank2()
    KDD = ixp.zerorank2()
    for i, j, k, l in itertools.product(R3, R3, R3, R3):
        gammaDD[i][j] += Jac[k][i] * Jac[l][j] * gS[k][l]
        KDD[i][j] += Jac[k][i] * Jac[l][j] * KS[k][l]
    BitoA.gammabarDD_hDD(gammaDD)
    BitoA.trK_AbarDD_aDD(gammaDD, KDD)
    BitoA.LambdabarU_lambdaU__exact_gammaDD(gammaDD)
    BitoA.cf_from_gammaDD(gammaDD)
    BitoA.betU_vetU(betaU, BU)
    return BitoA.cf, BitoA.hDD, BitoA.lambdaU, BitoA.aDD, BitoA.trK, alpha, BitoA.vetU, BitoA.betU
