"""BSSN gridfunctions and derived quantities in a reference-metric formulation.

Equations follow BSSN/BSSN_quantities.py (file:line cited per function). The
evolved variables are rescaled by the reference metric:
gammabar_ij = h_ij ReDD_ij + ghat_ij, Abar_ij = a_ij ReDD_ij,
Lambdabar^i = lambda^i ReU^i, beta^i = vet^i ReU^i, B^i = bet^i ReU^i.
Derivatives of gridfunctions are symbols named ``<gf>_dD<i>``,
``<gf>_dDD<ij>``, ``<gf>_dupD<i>`` (upwinded) that the finite-difference
code generator turns into stencils.
"""
import itertools
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import grid as gri
from .. import indexedexp as ixp
from .. import reference_metric as rfm

thismodule = "BSSN.BSSN_quantities"
par.initialize_param(par.glb_param("char", thismodule, "EvolvedConformalFactor_cf", "W"))
par.initialize_param(par.glb_param("bool", thismodule, "detgbarOverdetghat_equals_one", "True"))
par.initialize_param(par.glb_param("bool", thismodule, "LeaveRicciSymbolic", "False"))

R3 = range(3)


def _p3(n):
    return itertools.product(R3, repeat=n)


def declare_BSSN_gridfunctions_if_not_declared_already():
    """The 24 evolved BSSN gridfunctions (BSSN_quantities.py:43-70)."""
    global hDD, aDD, lambdaU, vetU, betU, trK, cf, alpha
    if any(gf.name == "hDD00" for gf in gri.glb_gridfcs_list):
        hDD = ixp.declarerank2("hDD", "sym01")
        aDD = ixp.declarerank2("aDD", "sym01")
        lambdaU = ixp.declarerank1("lambdaU")
        vetU = ixp.declarerank1("vetU")
        betU = ixp.declarerank1("betU")
        trK, cf, alpha = sp.symbols("trK cf alpha", real=True)
        return hDD, aDD, lambdaU, vetU, betU, trK, cf, alpha
    hDD = ixp.register_gridfunctions_for_single_rank2("EVOL", "hDD", "sym01")
    aDD = ixp.register_gridfunctions_for_single_rank2("EVOL", "aDD", "sym01")
    lambdaU = ixp.register_gridfunctions_for_single_rank1("EVOL", "lambdaU")
    vetU = ixp.register_gridfunctions_for_single_rank1("EVOL", "vetU")
    betU = ixp.register_gridfunctions_for_single_rank1("EVOL", "betU")
    trK, cf, alpha = gri.register_gridfunctions("EVOL", ["trK", "cf", "alpha"])
    return hDD, aDD, lambdaU, vetU, betU, trK, cf, alpha


def BSSN_basic_tensors():
    """gammabar_ij, Abar_ij, Lambdabar^i, beta^i, B^i (BSSN_quantities.py:76-103)."""
    global gammabarDD, AbarDD, LambdabarU, betaU, BU
    h, a, lam, vet, bet, _, _, _ = declare_BSSN_gridfunctions_if_not_declared_already()
    gammabarDD = [[h[i][j] * rfm.ReDD[i][j] + rfm.ghatDD[i][j] for j in R3] for i in R3]
    AbarDD = [[a[i][j] * rfm.ReDD[i][j] for j in R3] for i in R3]
    LambdabarU = [lam[i] * rfm.ReU[i] for i in R3]
    betaU = [vet[i] * rfm.ReU[i] for i in R3]
    BU = [bet[i] * rfm.ReU[i] for i in R3]


def gammabar__inverse_and_derivs():
    """gammabar^ij, gammabar_ij,k (centered and upwinded), gammabar_ij,kl and
    Gammabar^i_kl (BSSN_quantities.py:108-164)."""
    global gammabarUU, gammabarDD_dD, gammabarDD_dupD, gammabarDD_dDD, GammabarUDD
    h = declare_BSSN_gridfunctions_if_not_declared_already()[0]
    BSSN_basic_tensors()
    gammabarUU, _ = ixp.symm_matrix_inverter3x3(gammabarDD)
    h_dD = ixp.declarerank3("hDD_dD", "sym01")
    h_dupD = ixp.declarerank3("hDD_dupD", "sym01")
    h_dDD = ixp.declarerank4("hDD_dDD", "sym01_sym23")
    gammabarDD_dD = ixp.zerorank3()
    gammabarDD_dupD = ixp.zerorank3()
    for i, j, k in _p3(3):
        common = rfm.ghatDDdD[i][j][k] + h[i][j] * rfm.ReDDdD[i][j][k]
        gammabarDD_dD[i][j][k] = common + h_dD[i][j][k] * rfm.ReDD[i][j]
        gammabarDD_dupD[i][j][k] = common + h_dupD[i][j][k] * rfm.ReDD[i][j]
    gammabarDD_dDD = ixp.zerorank4()
    for i, j, k, l in _p3(4):
        gammabarDD_dDD[i][j][k][l] = (rfm.ghatDDdDD[i][j][k][l]
                                      + h_dDD[i][j][k][l] * rfm.ReDD[i][j]
                                      + h_dD[i][j][k] * rfm.ReDDdD[i][j][l]
                                      + h_dD[i][j][l] * rfm.ReDDdD[i][j][k]
                                      + h[i][j] * rfm.ReDDdDD[i][j][k][l])
    GammabarUDD = ixp.zerorank3()
    for i, k, l, m in _p3(4):
        GammabarUDD[i][k][l] += sp.Rational(1, 2) * gammabarUU[i][m] * (
            gammabarDD_dD[m][k][l] + gammabarDD_dD[m][l][k] - gammabarDD_dD[k][l][m])


def detgammabar_and_derivs():
    """det(gammabar) = det(ghat) and its derivatives (BSSN_quantities.py:168-203)."""
    global detgammabar, detgammabar_dD, detgammabar_dDD
    declare_BSSN_gridfunctions_if_not_declared_already()
    if par.parval_from_str(thismodule + "::detgbarOverdetghat_equals_one") == "False":
        raise NotImplementedError("detgbarOverdetghat_equals_one=False is not implemented")
    detgammabar = rfm.detgammahat
    detgammabar_dD = [rfm.detgammahatdD[i] for i in R3]
    detgammabar_dDD = [[rfm.detgammahatdDD[i][j] for j in R3] for i in R3]


def AbarUU_AbarUD_trAbar_AbarDD_dD():
    """Abar^ij, Abar^i_j, trace, Abar_ij,k (BSSN_quantities.py:209-255)."""
    global AbarUU, AbarUD, trAbar, AbarDD_dD, AbarDD_dupD
    a = declare_BSSN_gridfunctions_if_not_declared_already()[1]
    BSSN_basic_tensors()
    gammabar__inverse_and_derivs()
    AbarUU = ixp.zerorank2()
    for i, j, k, l in _p3(4):
        AbarUU[i][j] += gammabarUU[i][k] * gammabarUU[j][l] * AbarDD[k][l]
    AbarUD = ixp.zerorank2()
    for i, j, k in _p3(3):
        AbarUD[i][j] += gammabarUU[i][k] * AbarDD[k][j]
    trAbar = sp.sympify(0)
    for k, j in _p3(2):
        trAbar += gammabarUU[k][j] * AbarDD[j][k]
    a_dD = ixp.declarerank3("aDD_dD", "sym01")
    a_dupD = ixp.declarerank3("aDD_dupD", "sym01")
    AbarDD_dD = ixp.zerorank3()
    AbarDD_dupD = ixp.zerorank3()
    for i, j, k in _p3(3):
        AbarDD_dupD[i][j][k] = rfm.ReDDdD[i][j][k] * a[i][j] + rfm.ReDD[i][j] * a_dupD[i][j][k]
        AbarDD_dD[i][j][k] = rfm.ReDDdD[i][j][k] * a[i][j] + rfm.ReDD[i][j] * a_dD[i][j][k]


def RicciBar__gammabarDD_dHatD__DGammaUDD__DGammaU():
    """Conformal Ricci tensor in terms of hatted covariant derivatives of
    eps_ij = h_ij ReDD_ij (BSSN_quantities.py:260-425)."""
    global RbarDD, DGammaUDD, gammabarDD_dHatD, DGammaU
    h, _, lam = declare_BSSN_gridfunctions_if_not_declared_already()[0:3]
    gammabar__inverse_and_derivs()
    G = rfm.GammahatUDD
    GdD = rfm.GammahatUDDdD
    eps = [[h[i][j] * rfm.ReDD[i][j] for j in R3] for i in R3]
    h_dD = ixp.declarerank3("hDD_dD", "sym01")
    h_dDD = ixp.declarerank4("hDD_dDD", "sym01_sym23")
    eps_dD = ixp.zerorank3()
    for i, j, k in _p3(3):
        eps_dD[i][j][k] = h_dD[i][j][k] * rfm.ReDD[i][j] + h[i][j] * rfm.ReDDdD[i][j][k]
    eps_dDD = ixp.zerorank4()
    for i, j, k, l in _p3(4):
        eps_dDD[i][j][k][l] = (h_dDD[i][j][k][l] * rfm.ReDD[i][j] + h_dD[i][j][k] * rfm.ReDDdD[i][j][l]
                               + h_dD[i][j][l] * rfm.ReDDdD[i][j][k] + h[i][j] * rfm.ReDDdDD[i][j][k][l])
    # Dhat_l gammabar_ij
    gammabarDD_dHatD = ixp.zerorank3()
    for i, j, l in _p3(3):
        s = eps_dD[i][j][l]
        for m in R3:
            s += -G[m][i][l] * eps[m][j] - G[m][j][l] * eps[i][m]
        gammabarDD_dHatD[i][j][l] = s
    # partial_k of Dhat_l gammabar_ij
    dHatD_dD = ixp.zerorank4()
    for i, j, l, k in _p3(4):
        s = eps_dDD[i][j][l][k]
        for m in R3:
            s += (-GdD[m][i][l][k] * eps[m][j] - G[m][i][l] * eps_dD[m][j][k]
                  - GdD[m][j][l][k] * eps[i][m] - G[m][j][l] * eps_dD[i][m][k])
        dHatD_dD[i][j][l][k] = s
    # Dhat_k Dhat_l gammabar_ij
    dHatDD = ixp.zerorank4()
    for i, j, l, k in _p3(4):
        s = dHatD_dD[i][j][l][k]
        for m in R3:
            s += (-G[m][l][k] * gammabarDD_dHatD[i][j][m] - G[m][i][k] * gammabarDD_dHatD[m][j][l]
                  - G[m][j][k] * gammabarDD_dHatD[i][m][l])
        dHatDD[i][j][l][k] = s
    # Dhat_j Lambdabar^k
    lam_dD = ixp.declarerank2("lambdaU_dD", "nosym")
    LambarU_dHatD = ixp.zerorank2()
    for j, k in _p3(2):
        s = lam_dD[k][j] * rfm.ReU[k] + lam[k] * rfm.ReUdD[k][j]
        for m in R3:
            s += G[k][m][j] * lam[m] * rfm.ReU[m]
        LambarU_dHatD[k][j] = s
    DGammaUDD = ixp.zerorank3()
    for i, j, k in _p3(3):
        DGammaUDD[i][j][k] = GammabarUDD[i][j][k] - G[i][j][k]
    DGammaU = ixp.zerorank1()
    for i, j, k in _p3(3):
        DGammaU[i] += gammabarUU[j][k] * DGammaUDD[i][j][k]
    DGammaDDD = ixp.zerorank3()
    for i, j, k, m in _p3(4):
        DGammaDDD[i][j][k] += gammabarDD[i][m] * DGammaUDD[m][j][k]

    if par.parval_from_str(thismodule + "::LeaveRicciSymbolic") == "True":
        if not any(gf.name == "RbarDD00" for gf in gri.glb_gridfcs_list):
            ixp.register_gridfunctions_for_single_rank2("AUXEVOL", "RbarDD", "sym01")
        RbarDD = ixp.declarerank2("RbarDD", "sym01")
        return

    half = sp.Rational(1, 2)
    RbarDD = ixp.zerorank2()
    for i, j in _p3(2):
        s = sp.sympify(0)
        for k, l in _p3(2):
            s += -half * gammabarUU[k][l] * dHatDD[i][j][l][k]
        for k in R3:
            s += half * (gammabarDD[k][i] * LambarU_dHatD[k][j] + gammabarDD[k][j] * LambarU_dHatD[k][i])
        for k in R3:
            s += half * DGammaU[k] * (DGammaDDD[i][j][k] + DGammaDDD[j][i][k])
        for k, l, m in _p3(3):
            s += gammabarUU[k][l] * (DGammaUDD[m][k][i] * DGammaDDD[j][m][l]
                                     + DGammaUDD[m][k][j] * DGammaDDD[i][m][l]
                                     + DGammaUDD[m][i][k] * DGammaDDD[m][j][l])
        RbarDD[i][j] = s


def betaU_derivs():
    """beta^i_,j (centered and upwinded) and beta^i_,jk (BSSN_quantities.py:431-453)."""
    global betaU_dD, betaU_dupD, betaU_dDD
    vet = declare_BSSN_gridfunctions_if_not_declared_already()[3]
    vet_dD = ixp.declarerank2("vetU_dD", "nosym")
    vet_dupD = ixp.declarerank2("vetU_dupD", "nosym")
    vet_dDD = ixp.declarerank3("vetU_dDD", "sym12")
    betaU_dD = ixp.zerorank2()
    betaU_dupD = ixp.zerorank2()
    betaU_dDD = ixp.zerorank3()
    for i, j in _p3(2):
        betaU_dD[i][j] = vet_dD[i][j] * rfm.ReU[i] + vet[i] * rfm.ReUdD[i][j]
        betaU_dupD[i][j] = vet_dupD[i][j] * rfm.ReU[i] + vet[i] * rfm.ReUdD[i][j]
        for k in R3:
            betaU_dDD[i][j][k] = (vet_dDD[i][j][k] * rfm.ReU[i] + vet_dD[i][j] * rfm.ReUdD[i][k]
                                  + vet_dD[i][k] * rfm.ReUdD[i][j] + vet[i] * rfm.ReUdDD[i][j][k])


def phi_and_derivs():
    """phi derivatives from the evolved conformal factor cf (W, phi or chi),
    exp(-4 phi), and Dbar_i Dbar_j phi (BSSN_quantities.py:459-534)."""
    global phi_dD, phi_dupD, phi_dDD, exp_m4phi, phi_dBarD, phi_dBarDD
    cf_ = declare_BSSN_gridfunctions_if_not_declared_already()[6]
    gammabar__inverse_and_derivs()
    cf_dD = ixp.declarerank1("cf_dD")
    cf_dupD = ixp.declarerank1("cf_dupD")
    cf_dDD = ixp.declarerank2("cf_dDD", "sym01")
    choice = par.parval_from_str(thismodule + "::EvolvedConformalFactor_cf")
    phi_dD = ixp.zerorank1()
    phi_dupD = ixp.zerorank1()
    phi_dDD = ixp.zerorank2()
    if choice == "phi":
        for i in R3:
            phi_dD[i] = cf_dD[i]
            phi_dupD[i] = cf_dupD[i]
            for j in R3:
                phi_dDD[i][j] = cf_dDD[i][j]
        exp_m4phi = sp.exp(-4 * cf_)
    elif choice in ("W", "chi"):
        # W = e^{-2 phi}  -> phi_,i = -W_,i/(2W);  chi = e^{-4 phi} -> phi_,i = -chi_,i/(4 chi)
        n = 2 if choice == "W" else 4
        for i in R3:
            phi_dD[i] = -cf_dD[i] / (n * cf_)
            phi_dupD[i] = -cf_dupD[i] / (n * cf_)
            for j in R3:
                phi_dDD[i][j] = (-cf_dDD[i][j] + cf_dD[i] * cf_dD[j] / cf_) / (n * cf_)
        exp_m4phi = cf_ * cf_ if choice == "W" else cf_
    else:
        raise ValueError("EvolvedConformalFactor_cf == " + choice + " unsupported")
    phi_dBarD = phi_dD
    phi_dBarDD = ixp.zerorank2()
    for i, j in _p3(2):
        s = phi_dDD[i][j]
        for k in R3:
            s += -GammabarUDD[k][i][j] * phi_dD[k]
        phi_dBarDD[i][j] = s
