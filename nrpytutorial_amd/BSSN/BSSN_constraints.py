"""Hamiltonian and momentum constraints (BSSN/BSSN_constraints.py:19-126).

H  = 2/3 K^2 - Abar_ij Abar^ij + e^{-4phi} (Rbar - 8 (gammabar^ij phi_,i phi_,j
     + gammabar^ij Dbar_i Dbar_j phi))                         (:49-80)
M^i = e^{-4phi}/ReU^i [6 Abar^ij phi_,j - 2/3 gammabar^ij K_,j
     + gammabar^ik gammabar^jl Dbar_j Abar_kl]                 (:93-126)
"""
import itertools
import sympy as sp
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from . import BSSN_quantities as Bq

R3 = range(3)


def BSSN_constraints(add_T4UUmunu_source_terms=False):
    global H, MU
    if add_T4UUmunu_source_terms:
        raise NotImplementedError("matter source terms are not part of the vacuum hot path")
    rfm.reference_metric()
    Bq.declare_BSSN_gridfunctions_if_not_declared_already()
    Bq.BSSN_basic_tensors()
    Bq.gammabar__inverse_and_derivs()
    Bq.AbarUU_AbarUD_trAbar_AbarDD_dD()
    Bq.RicciBar__gammabarDD_dHatD__DGammaUDD__DGammaU()
    Bq.phi_and_derivs()
    gUU = Bq.gammabarUU
    H = sp.Rational(2, 3) * Bq.trK ** 2
    for i, j in itertools.product(R3, R3):
        H += -Bq.AbarDD[i][j] * Bq.AbarUU[i][j]
    Rbar = sp.sympify(0)
    dphi2 = sp.sympify(0)
    ddphi = sp.sympify(0)
    for i, j in itertools.product(R3, R3):
        Rbar += gUU[i][j] * Bq.RbarDD[i][j]
        dphi2 += gUU[i][j] * Bq.phi_dBarD[i] * Bq.phi_dBarD[j]
        ddphi += gUU[i][j] * Bq.phi_dBarDD[i][j]
    H += Bq.exp_m4phi * (Rbar - 8 * (dphi2 + ddphi))

    MU = ixp.zerorank1()
    for i, j in itertools.product(R3, R3):
        MU[i] += 6 * Bq.AbarUU[i][j] * Bq.phi_dD[j]
    trK_dD = ixp.declarerank1("trK_dD")
    for i, j in itertools.product(R3, R3):
        MU[i] += -sp.Rational(2, 3) * gUU[i][j] * trK_dD[j]
    AbarDD_dBarD = ixp.zerorank3()
    for i, j, k in itertools.product(R3, R3, R3):
        s = Bq.AbarDD_dD[i][j][k]
        for l in R3:
            s += -Bq.GammabarUDD[l][k][i] * Bq.AbarDD[l][j] - Bq.GammabarUDD[l][k][j] * Bq.AbarDD[i][l]
        AbarDD_dBarD[i][j][k] = s
    for i, j, k, l in itertools.product(R3, R3, R3, R3):
        MU[i] += gUU[i][k] * gUU[j][l] * AbarDD_dBarD[k][l][j]
    for i in R3:
        MU[i] *= Bq.exp_m4phi / rfm.ReU[i]
