"""Vacuum Maxwell System II in a curvilinear reference metric, vectors rescaled
by the reference metric, A^i = a^i ReU^i, E^i = e^i ReU^i
(Maxwell/VacuumMaxwell_Flat_Evol_Curvilinear_rescaled.py:11-196).

With hatted covariant derivatives of the flat reference metric:
  da^i/dt  = -e^i - ghat^ij d_j psi / ReU^i
  dE^i/dt  = ghat^ij d_j Gamma - ghat^jk (Dhat_j Dhat_k A)^i,   e_rhs^i = E_rhs^i / ReU^i
  dGamma/dt = -ghat^ij (d_i d_j psi - Gammahat^k_ji d_k psi)
  dpsi/dt  = -Gamma
  G = Gamma - Dhat_i A^i,  C = Dhat_i E^i
where (Dhat_j Dhat_k A)^i expands into second partials of A^i, first partials
times Gammahat, derivatives of Gammahat, and Gammahat x Gammahat terms.
AU_Cart, EU_Cart are A^i, E^i in the Cartesian basis.
"""
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import grid as gri
from .. import indexedexp as ixp
from .. import reference_metric as rfm

R3 = range(3)


def VacuumMaxwellRHSs_rescaled():
    global erhsU, arhsU, psi_rhs, Gamma_rhs, C, G, EU_Cart, AU_Cart
    par.set_parval_from_str("grid::DIM", 3)
    rfm.reference_metric()
    eU = ixp.register_gridfunctions_for_single_rank1("EVOL", "eU")
    eU_dD = ixp.declarerank2("eU_dD", "nosym")
    aU = ixp.register_gridfunctions_for_single_rank1("EVOL", "aU")
    aU_dD = ixp.declarerank2("aU_dD", "nosym")
    aU_dDD = ixp.declarerank3("aU_dDD", "sym12")
    gri.register_gridfunctions("EVOL", ["psi"])
    Gamma = gri.register_gridfunctions("EVOL", ["Gamma"])
    psi_dD = ixp.declarerank1("psi_dD")
    Gamma_dD = ixp.declarerank1("Gamma_dD")
    psi_dDD = ixp.declarerank2("psi_dDD", "sym01")
    gUU, Ghat, GhatdD = rfm.ghatUU, rfm.GammahatUDD, rfm.GammahatUDDdD
    ReU, ReUdD, ReUdDD = rfm.ReU, rfm.ReUdD, rfm.ReUdDD

    arhsU = [-eU[i] - sum(gUU[i][j] * psi_dD[j] for j in R3) / ReU[i] for i in R3]
    # Unrescaled A^i and its first and second partial derivatives (product rule).
    AU = [aU[i] * ReU[i] for i in R3]
    AU_dD = [[aU_dD[i][j] * ReU[i] + aU[i] * ReUdD[i][j] for j in R3] for i in R3]
    AU_dDD = [[[aU_dDD[i][j][k] * ReU[i] + aU_dD[i][j] * ReUdD[i][k] + aU_dD[i][k] * ReUdD[i][j]
                + aU[i] * ReUdDD[i][j][k] for k in R3] for j in R3] for i in R3]
    # (Dhat_j Dhat_k A)^i, written as three pieces: second partials, terms linear in
    # Gammahat and its derivative, and terms quadratic in Gammahat.
    DDA = [[[AU_dDD[i][k][j]
             + sum(GhatdD[i][m][k][j] * AU[m] + Ghat[i][m][k] * AU_dD[m][j] + Ghat[i][m][j] * AU_dD[m][k]
                   - Ghat[m][k][j] * AU_dD[i][m] for m in R3)
             + sum((Ghat[i][d][j] * Ghat[d][m][k] - Ghat[d][k][j] * Ghat[i][m][d]) * AU[m]
                   for m in R3 for d in R3)
             for k in R3] for j in R3] for i in R3]
    ErhsU = [sum(gUU[i][j] * Gamma_dD[j] for j in R3) - sum(gUU[j][k] * DDA[i][j][k] for j in R3 for k in R3)
             for i in R3]
    erhsU = [ErhsU[i] / ReU[i] for i in R3]
    Gamma_rhs = sp.sympify(0)
    for i in R3:
        for j in R3:
            Gamma_rhs += -gUU[i][j] * psi_dDD[i][j] + sum(gUU[i][j] * Ghat[k][j][i] * psi_dD[k] for k in R3)
    psi_rhs = -Gamma
    G = Gamma - sum(AU_dD[i][i] for i in R3) + sum(Ghat[i][j][i] * AU[j] for i in R3 for j in R3)
    EU = [eU[i] * ReU[i] for i in R3]
    EU_dD = [[eU_dD[i][j] * ReU[i] + eU[i] * ReUdD[i][j] for j in R3] for i in R3]
    C = sum(EU_dD[i][i] for i in R3) + sum(Ghat[i][j][i] * EU[j] for i in R3 for j in R3)
    J = [[sp.diff(rfm.xx_to_Cart[i], rfm.xx[j]) for j in R3] for i in R3]
    AU_Cart = [sum(J[i][j] * AU[j] for j in R3) for i in R3]
    EU_Cart = [sum(J[i][j] * EU[j] for j in R3) for i in R3]
