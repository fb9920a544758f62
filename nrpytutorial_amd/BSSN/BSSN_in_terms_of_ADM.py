"""BSSN variables from ADM variables (BSSN/BSSN_in_terms_of_ADM.py:28-171).

gammabar_ij = (det ghat / det gamma)^{1/3} gamma_ij,  h_ij = (gammabar_ij - ghat_ij)/ReDD_ij
Abar_ij = (det ghat / det gamma)^{1/3} (K_ij - gamma_ij K/3),  a_ij = Abar_ij / ReDD_ij
Lambdabar^i = gammabar^jk (Gammabar^i_jk - Gammahat^i_jk) from exact derivatives
W = (det gamma / det gammabar)^{-1/6}   (phi and chi variants too)
vet^i = beta^i / ReU^i,  bet^i = B^i / ReU^i
"""
import itertools
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from . import BSSN_quantities as Bq  # registers EvolvedConformalFactor_cf

R3 = range(3)


def _need_rfm():
    if not rfm.have_already_called_reference_metric_function:
        raise RuntimeError("call reference_metric() first")


def gammabarDD_hDD(gammaDD):
    global gammabarDD, hDD
    if gammaDD is None:
        gammaDD = ixp.declarerank2("gammaDD", "sym01")
    _need_rfm()
    _, gammaDET = ixp.symm_matrix_inverter3x3(gammaDD)
    gammabarDD = [[(rfm.detgammahat / gammaDET) ** sp.Rational(1, 3) * gammaDD[i][j] for j in R3] for i in R3]
    hDD = [[(gammabarDD[i][j] - rfm.ghatDD[i][j]) / rfm.ReDD[i][j] for j in R3] for i in R3]


def trK_AbarDD_aDD(gammaDD, KDD):
    global trK, AbarDD, aDD
    if gammaDD is None:
        gammaDD = ixp.declarerank2("gammaDD", "sym01")
    if KDD is None:
        KDD = ixp.declarerank2("KDD", "sym01")
    _need_rfm()
    gammaUU, gammaDET = ixp.symm_matrix_inverter3x3(gammaDD)
    trK = sum(gammaUU[i][j] * KDD[i][j] for i, j in itertools.product(R3, R3))
    AbarDD = [[(rfm.detgammahat / gammaDET) ** sp.Rational(1, 3) * (KDD[i][j] - sp.Rational(1, 3) * gammaDD[i][j] * trK)
               for j in R3] for i in R3]
    aDD = [[AbarDD[i][j] / rfm.ReDD[i][j] for j in R3] for i in R3]


def LambdabarU_lambdaU__exact_gammaDD(gammaDD):
    global LambdabarU, lambdaU
    if gammaDD is None:
        gammaDD = ixp.declarerank2("gammaDD", "sym01")
    gammabarDD_hDD(gammaDD)
    gbarUU, _ = ixp.symm_matrix_inverter3x3(gammabarDD)
    GbarUDD = ixp.zerorank3()
    for i, j, k, l in itertools.product(R3, R3, R3, R3):
        GbarUDD[i][j][k] += sp.Rational(1, 2) * gbarUU[i][l] * (
            sp.diff(gammabarDD[l][j], rfm.xx[k]) + sp.diff(gammabarDD[l][k], rfm.xx[j])
            - sp.diff(gammabarDD[j][k], rfm.xx[l]))
    LambdabarU = ixp.zerorank1()
    for i, j, k in itertools.product(R3, R3, R3):
        LambdabarU[i] += gbarUU[j][k] * (GbarUDD[i][j][k] - rfm.GammahatUDD[i][j][k])
    LambdabarU = [sp.sympify(L).doit() for L in LambdabarU]
    lambdaU = [LambdabarU[i] / rfm.ReU[i] for i in R3]


def cf_from_gammaDD(gammaDD):
    global cf
    if gammaDD is None:
        gammaDD = ixp.declarerank2("gammaDD", "sym01")
    gammabarDD_hDD(gammaDD)
    _, gbarDET = ixp.symm_matrix_inverter3x3(gammabarDD)
    _, gDET = ixp.symm_matrix_inverter3x3(gammaDD)
    choice = par.parval_from_str("EvolvedConformalFactor_cf")
    if choice == "phi":
        cf = sp.Rational(1, 12) * sp.log(gDET / gbarDET)
    elif choice == "chi":
        cf = (gDET / gbarDET) ** (-sp.Rational(1, 3))
    elif choice == "W":
        cf = (gDET / gbarDET) ** (-sp.Rational(1, 6))
    else:
        raise ValueError("EvolvedConformalFactor_cf = " + choice + " unknown")


def betU_vetU(betaU, BU):
    global vetU, betU
    if betaU is None:
        betaU = ixp.declarerank1("betaU")
    if BU is None:
        BU = ixp.declarerank1("BU")
    _need_rfm()
    vetU = [betaU[i] / rfm.ReU[i] for i in R3]
    betU = [BU[i] / rfm.ReU[i] for i in R3]
