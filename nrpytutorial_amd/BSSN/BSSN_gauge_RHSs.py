"""Gauge right-hand sides: lapse and shift evolution (BSSN/BSSN_gauge_RHSs.py).

Lapse options (:54-91): OnePlusLog, HarmonicSlicing, Frozen, OnePlusLogAlt.
Shift options (:96-235): Frozen, GammaDriving2ndOrder_NoCovariant,
GammaDriving2ndOrder_Covariant[__Hatted], GammaDriving1stOrder_Covariant[__Hatted],
NonAdvectingGammaDriving. Outputs are rescaled by ReU (:286-293).
"""
import itertools
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from . import BSSN_quantities as Bq
from . import BSSN_RHSs as Brhs

thismodule = "BSSN.BSSN_gauge_RHSs"
par.initialize_param(par.glb_param("char", thismodule, "LapseEvolutionOption", "OnePlusLog"))
par.initialize_param(par.glb_param("char", thismodule, "ShiftEvolutionOption", "GammaDriving2ndOrder_Covariant"))
R3 = range(3)

SHIFT_OPTIONS = ("Frozen", "GammaDriving2ndOrder_NoCovariant", "GammaDriving2ndOrder_Covariant",
                 "GammaDriving2ndOrder_Covariant__Hatted", "GammaDriving1stOrder_Covariant",
                 "GammaDriving1stOrder_Covariant__Hatted", "NonAdvectingGammaDriving")


def BSSN_gauge_RHSs():
    global alpha_rhs, vet_rhsU, bet_rhsU
    rfm.reference_metric()
    Bq.BSSN_basic_tensors()
    Bq.betaU_derivs()
    if not Brhs.have_already_called_BSSN_RHSs_function:
        raise RuntimeError("BSSN_gauge_RHSs() requires BSSN_RHSs() to be called first")
    cf, trK, alpha, betaU = Bq.cf, Bq.trK, Bq.alpha, Bq.betaU

    lapse = par.parval_from_str(thismodule + "::LapseEvolutionOption")
    if lapse in ("OnePlusLog", "OnePlusLogAlt"):
        alpha_rhs = -2 * alpha * trK if lapse == "OnePlusLog" else -alpha * (1 - alpha) * trK
        alpha_dupD = ixp.declarerank1("alpha_dupD")
        for i in R3:
            alpha_rhs += betaU[i] * alpha_dupD[i]
    elif lapse == "HarmonicSlicing":
        choice = par.parval_from_str("BSSN.BSSN_quantities::EvolvedConformalFactor_cf")
        if choice == "W":
            alpha_rhs = -3 * cf ** (-4) * Brhs.cf_rhs
        elif choice == "phi":
            alpha_rhs = 6 * sp.exp(6 * cf) * Brhs.cf_rhs
        else:
            raise ValueError("HarmonicSlicing requires cf = W or phi")
    elif lapse == "Frozen":
        alpha_rhs = sp.sympify(0)
    else:
        raise ValueError("LapseEvolutionOption == " + lapse + " not supported")

    shift = par.parval_from_str(thismodule + "::ShiftEvolutionOption")
    if shift not in SHIFT_OPTIONS:
        raise ValueError("ShiftEvolutionOption == " + shift + " unsupported")
    BU, betU, betaU_dupD = Bq.BU, Bq.betU, Bq.betaU_dupD
    beta_rhsU = ixp.zerorank1()
    B_rhsU = ixp.zerorank1()

    def BU_dupD_of():
        betU_dupD = ixp.declarerank2("betU_dupD", "nosym")
        return [[betU_dupD[i][j] * rfm.ReU[i] + betU[i] * rfm.ReUdD[i][j] for j in R3] for i in R3]

    if shift == "GammaDriving2ndOrder_NoCovariant":
        for i in R3:
            beta_rhsU[i] += BU[i]
            for j in R3:
                beta_rhsU[i] += betaU[j] * betaU_dupD[i][j]
        eta = par.Cparameters("REAL", thismodule, ["eta"], 2.0)
        BU_dupD = BU_dupD_of()
        Lam0 = [Brhs.Lambdabar_rhsU[i] for i in R3]
        for i, j in itertools.product(R3, R3):
            Lam0[j] += -betaU[i] * Brhs.LambdabarU_dupD[j][i]
        for i in R3:
            B_rhsU[i] += sp.Rational(3, 4) * Lam0[i] - eta * BU[i]
            for j in R3:
                B_rhsU[i] += betaU[j] * BU_dupD[i][j]

    if "GammaDriving2ndOrder_Covariant" in shift:
        Bq.gammabar__inverse_and_derivs()
        Conn = rfm.GammahatUDD if shift.endswith("__Hatted") else Bq.GammabarUDD
        for i, j in itertools.product(R3, R3):
            beta_rhsU[i] += betaU[j] * betaU_dupD[i][j]
        for i, j, m in itertools.product(R3, R3, R3):
            beta_rhsU[i] += betaU[j] * Conn[i][m][j] * betaU[m]
        for i in R3:
            beta_rhsU[i] += BU[i]
        BU_dupD = BU_dupD_of()
        for i, j in itertools.product(R3, R3):
            B_rhsU[i] += betaU[j] * BU_dupD[i][j]
        for i, j, m in itertools.product(R3, R3, R3):
            B_rhsU[i] += betaU[j] * Conn[i][m][j] * BU[m]
        for i in R3:
            B_rhsU[i] += sp.Rational(3, 4) * Brhs.Lambdabar_rhsU[i]
        for i, j in itertools.product(R3, R3):
            B_rhsU[i] += -sp.Rational(3, 4) * betaU[j] * Brhs.LambdabarU_dupD[i][j]
        for i, j, m in itertools.product(R3, R3, R3):
            B_rhsU[i] += -sp.Rational(3, 4) * betaU[j] * Conn[i][m][j] * Bq.LambdabarU[m]
        eta = par.Cparameters("REAL", thismodule, ["eta"], 2.0)
        for i in R3:
            B_rhsU[i] += -eta * BU[i]

    if "GammaDriving1stOrder_Covariant" in shift:
        B_rhsU = ixp.zerorank1()
        Bq.gammabar__inverse_and_derivs()
        Conn = rfm.GammahatUDD if shift.endswith("__Hatted") else Bq.GammabarUDD
        for i, j in itertools.product(R3, R3):
            beta_rhsU[i] += betaU[j] * betaU_dupD[i][j]
        for i, j, m in itertools.product(R3, R3, R3):
            beta_rhsU[i] += betaU[j] * Conn[i][m][j] * betaU[m]
        eta = par.Cparameters("REAL", thismodule, ["eta"], 2.0)
        for i in R3:
            beta_rhsU[i] += sp.Rational(3, 4) * Bq.LambdabarU[i] - eta * betaU[i]

    if shift == "NonAdvectingGammaDriving":
        eta = par.Cparameters("REAL", thismodule, ["eta"], 2.0)
        for i in R3:
            beta_rhsU[i] += BU[i]
            B_rhsU[i] += sp.Rational(3, 4) * Brhs.Lambdabar_rhsU[i] - eta * BU[i]

    vet_rhsU = [beta_rhsU[i] / rfm.ReU[i] for i in R3]
    bet_rhsU = [B_rhsU[i] / rfm.ReU[i] for i in R3]
