"""Right-hand sides of the BSSN evolution equations (non-gauge part).

Equations follow BSSN/BSSN_RHSs.py: gammabar_ij (:66-100), Abar_ij (:102-169),
cf (:171-190), K (:192-212), Lambdabar^i (:214-316), then rescaling to the
evolved h_ij, a_ij, lambda^i (:318-328). Advection terms use the upwinded
derivative symbols ``*_dupD``.
"""
import itertools
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from . import BSSN_quantities as Bq

have_already_called_BSSN_RHSs_function = False
R3 = range(3)


def _p3(n):
    return itertools.product(R3, repeat=n)


def BSSN_RHSs():
    global have_already_called_BSSN_RHSs_function
    global cf_rhs, trK_rhs, LambdabarU_dupD, Lambdabar_rhsU, h_rhsDD, a_rhsDD, lambda_rhsU
    rfm.reference_metric()
    have_already_called_BSSN_RHSs_function = True

    Bq.BSSN_basic_tensors()
    gammabarDD, AbarDD, LambdabarU, betaU = Bq.gammabarDD, Bq.AbarDD, Bq.LambdabarU, Bq.betaU
    trK, alpha, cf, lambdaU = Bq.trK, Bq.alpha, Bq.cf, Bq.lambdaU
    Bq.betaU_derivs()
    betaU_dD, betaU_dDD = Bq.betaU_dD, Bq.betaU_dDD
    Bq.gammabar__inverse_and_derivs()
    gammabarDD_dupD = Bq.gammabarDD_dupD
    Bq.AbarUU_AbarUD_trAbar_AbarDD_dD()
    trAbar, AbarDD_dupD, gammabarUU, AbarUD, AbarUU = Bq.trAbar, Bq.AbarDD_dupD, Bq.gammabarUU, Bq.AbarUD, Bq.AbarUU
    Bq.detgammabar_and_derivs()
    detgammabar, detgammabar_dD, detgammabar_dDD = Bq.detgammabar, Bq.detgammabar_dD, Bq.detgammabar_dDD

    # Dbar_k beta^k = beta^k_,k + beta^k (det gammabar)_,k / (2 det gammabar)
    Dbarbetacontraction = sp.sympify(0)
    for k in R3:
        Dbarbetacontraction += betaU_dD[k][k] + betaU[k] * detgammabar_dD[k] / (2 * detgammabar)

    # partial_t gammabar_ij: Lie derivative along beta, trace term, -2 alpha Abar_ij
    gammabar_rhsDD = ixp.zerorank2()
    for i, j in _p3(2):
        s = sp.sympify(0)
        for k in R3:
            s += (betaU[k] * gammabarDD_dupD[i][j][k] + betaU_dD[k][i] * gammabarDD[k][j]
                  + betaU_dD[k][j] * gammabarDD[i][k])
        s += sp.Rational(2, 3) * gammabarDD[i][j] * (alpha * trAbar - Dbarbetacontraction)
        s += -2 * alpha * AbarDD[i][j]
        gammabar_rhsDD[i][j] = s

    # partial_t Abar_ij
    Abar_rhsDD = ixp.zerorank2()
    for i, j in _p3(2):
        s = sp.sympify(0)
        for k in R3:
            s += (betaU[k] * AbarDD_dupD[i][j][k] + betaU_dD[k][i] * AbarDD[k][j]
                  + betaU_dD[k][j] * AbarDD[i][k])
        s += -sp.Rational(2, 3) * AbarDD[i][j] * Dbarbetacontraction + alpha * AbarDD[i][j] * trK
        for k in R3:
            s += -2 * alpha * AbarDD[i][k] * AbarUD[k][j]
        Abar_rhsDD[i][j] = s

    Bq.phi_and_derivs()
    phi_dD, phi_dupD, exp_m4phi = Bq.phi_dD, Bq.phi_dupD, Bq.exp_m4phi
    phi_dBarD, phi_dBarDD = Bq.phi_dBarD, Bq.phi_dBarDD
    Bq.RicciBar__gammabarDD_dHatD__DGammaUDD__DGammaU()
    RbarDD = Bq.RbarDD

    alpha_dD = ixp.declarerank1("alpha_dD")
    alpha_dDD = ixp.declarerank2("alpha_dDD", "sym01")
    GammabarUDD = Bq.GammabarUDD
    alpha_dBarD = alpha_dD
    alpha_dBarDD = ixp.zerorank2()
    for i, j in _p3(2):
        s = alpha_dDD[i][j]
        for k in R3:
            s += -GammabarUDD[k][i][j] * alpha_dD[k]
        alpha_dBarDD[i][j] = s

    # trace-free part of [-2 alpha Dbar_i Dbar_j phi + ... + alpha Rbar_ij], times e^{-4phi}
    curlyDD = ixp.zerorank2()
    for i, j in _p3(2):
        curlyDD[i][j] = (-2 * alpha * phi_dBarDD[i][j] + 4 * alpha * phi_dBarD[i] * phi_dBarD[j]
                         + 2 * alpha_dBarD[i] * phi_dBarD[j] + 2 * alpha_dBarD[j] * phi_dBarD[i]
                         - alpha_dBarDD[i][j] + alpha * RbarDD[i][j])
    curly_trace = sp.sympify(0)
    for i, j in _p3(2):
        curly_trace += gammabarUU[i][j] * curlyDD[i][j]
    for i, j in _p3(2):
        Abar_rhsDD[i][j] += exp_m4phi * (curlyDD[i][j] - sp.Rational(1, 3) * gammabarDD[i][j] * curly_trace)

    # partial_t cf
    phi_rhs = sp.Rational(1, 6) * (Dbarbetacontraction - alpha * trK)
    for k in R3:
        phi_rhs += betaU[k] * phi_dupD[k]
    choice = par.parval_from_str("BSSN.BSSN_quantities::EvolvedConformalFactor_cf")
    if choice == "phi":
        cf_rhs = phi_rhs
    elif choice == "W":
        cf_rhs = phi_rhs * (-2 * cf)
    elif choice == "chi":
        cf_rhs = phi_rhs * (-4 * cf)
    else:
        raise ValueError("EvolvedConformalFactor_cf == " + choice + " unsupported")

    # partial_t K
    trK_rhs = sp.Rational(1, 3) * alpha * trK * trK
    trK_dupD = ixp.declarerank1("trK_dupD")
    for i in R3:
        trK_rhs += betaU[i] * trK_dupD[i]
    for i, j in _p3(2):
        trK_rhs += -exp_m4phi * gammabarUU[i][j] * (alpha_dBarDD[i][j] + 2 * alpha_dBarD[j] * phi_dBarD[i])
    for i, j in _p3(2):
        trK_rhs += alpha * AbarDD[i][j] * AbarUU[i][j]

    # partial_t Lambdabar^i
    lambdaU_dupD = ixp.declarerank2("lambdaU_dupD", "nosym")
    LambdabarU_dupD = ixp.zerorank2()
    for i, j in _p3(2):
        LambdabarU_dupD[i][j] = lambdaU_dupD[i][j] * rfm.ReU[i] + lambdaU[i] * rfm.ReUdD[i][j]
    G, GdD = rfm.GammahatUDD, rfm.GammahatUDDdD
    Lambdabar_rhsU = ixp.zerorank1()
    for i, k in _p3(2):
        Lambdabar_rhsU[i] += betaU[k] * LambdabarU_dupD[i][k] - betaU_dD[i][k] * LambdabarU[k]
    # gammabar^jk Dhat_j Dhat_k beta^i, expanded
    for i, j, k in _p3(3):
        term = betaU_dDD[i][k][j]
        for m in R3:
            term += (GdD[i][m][k][j] * betaU[m] + G[i][m][k] * betaU_dD[m][j]
                     + G[i][m][j] * betaU_dD[m][k] - G[m][k][j] * betaU_dD[i][m])
        for m, d in _p3(2):
            term += (G[i][d][j] * G[d][m][k] - G[d][k][j] * G[i][m][d]) * betaU[m]
        Lambdabar_rhsU[i] += gammabarUU[j][k] * term
    DGammaU = Bq.DGammaU
    for i in R3:
        Lambdabar_rhsU[i] += sp.Rational(2, 3) * DGammaU[i] * Dbarbetacontraction
    Dbarbeta_dBarD = ixp.zerorank1()
    for k, m in _p3(2):
        Dbarbeta_dBarD[m] += (betaU_dDD[k][k][m]
                              + (betaU_dD[k][m] * detgammabar_dD[k] + betaU[k] * detgammabar_dDD[k][m]) / (2 * detgammabar)
                              - betaU[k] * detgammabar_dD[k] * detgammabar_dD[m] / (2 * detgammabar * detgammabar))
    for i, m in _p3(2):
        Lambdabar_rhsU[i] += sp.Rational(1, 3) * gammabarUU[i][m] * Dbarbeta_dBarD[m]
    for i, j in _p3(2):
        Lambdabar_rhsU[i] += -2 * AbarUU[i][j] * (alpha_dD[j] - 6 * alpha * phi_dD[j])
    DGammaUDD = Bq.DGammaUDD
    for i, j, k in _p3(3):
        Lambdabar_rhsU[i] += 2 * alpha * AbarUU[j][k] * DGammaUDD[i][j][k]
    trK_dD = ixp.declarerank1("trK_dD")
    for i, j in _p3(2):
        Lambdabar_rhsU[i] += -sp.Rational(4, 3) * alpha * gammabarUU[i][j] * trK_dD[j]

    h_rhsDD = [[gammabar_rhsDD[i][j] / rfm.ReDD[i][j] for j in R3] for i in R3]
    a_rhsDD = [[Abar_rhsDD[i][j] / rfm.ReDD[i][j] for j in R3] for i in R3]
    lambda_rhsU = [Lambdabar_rhsU[i] / rfm.ReU[i] for i in R3]
