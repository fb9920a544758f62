"""Brill-Lindquist two-black-hole initial data (BSSN/BrillLindquist.py:36-82).

psi = 1 + m1/(2|x-x1|) + m2/(2|x-x2|),  gamma_ij = psi^4 delta_ij, K_ij = 0,
alpha = psi^{-2}, beta = B = 0; converted to rescaled BSSN variables.
"""
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import indexedexp as ixp
from . import ADM_Exact_Spherical_or_Cartesian_to_BSSNCurvilinear as AtoB

thismodule = "BSSN.BrillLindquist"
BH1_posn_x, BH1_posn_y, BH1_posn_z = par.Cparameters("REAL", thismodule, ["BH1_posn_x", "BH1_posn_y", "BH1_posn_z"],
                                                     [0.0, 0.0, +0.5])
BH1_mass = par.Cparameters("REAL", thismodule, ["BH1_mass"], 1.0)
BH2_posn_x, BH2_posn_y, BH2_posn_z = par.Cparameters("REAL", thismodule, ["BH2_posn_x", "BH2_posn_y", "BH2_posn_z"],
                                                     [0.0, 0.0, -0.5])
BH2_mass = par.Cparameters("REAL", thismodule, ["BH2_mass"], 1.0)


def BrillLindquist(ComputeADMGlobalsOnly=False):
    global Cartxyz, gammaCartDD, KCartDD, alphaCart, betaCartU, BCartU
    global cf, hDD, lambdaU, aDD, trK, alpha, vetU, betU
    par.set_parval_from_str("grid::DIM", 3)
    Cartxyz = ixp.declarerank1("Cartxyz")
    x = Cartxyz
    psi = (sp.sympify(1)
           + BH1_mass / (2 * sp.sqrt((x[0] - BH1_posn_x) ** 2 + (x[1] - BH1_posn_y) ** 2 + (x[2] - BH1_posn_z) ** 2))
           + BH2_mass / (2 * sp.sqrt((x[0] - BH2_posn_x) ** 2 + (x[1] - BH2_posn_y) ** 2 + (x[2] - BH2_posn_z) ** 2)))
    gammaCartDD = ixp.zerorank2()
    for i in range(3):
        gammaCartDD[i][i] = psi ** 4
    KCartDD = ixp.zerorank2()
    alphaCart = 1 / psi ** 2
    betaCartU = ixp.zerorank1()
    BCartU = ixp.zerorank1()
    if ComputeADMGlobalsOnly:
        return
    cf, hDD, lambdaU, aDD, trK, alpha, vetU, betU = \
        AtoB.Convert_Spherical_or_Cartesian_ADM_to_BSSN_curvilinear("Cartesian", Cartxyz, gammaCartDD, KCartDD,
                                                                    alphaCart, betaCartU, BCartU)
