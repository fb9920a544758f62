"""Vacuum Maxwell in flat space, Cartesian coordinates
(Maxwell/VacuumMaxwell_Flat_Evol_Cartesian.py:54-112; Knapp, Walker & Baumgarte,
gr-qc/0201051, eqs. 10-15).

Evolved: A^i, E^i, psi, and for System II the auxiliary Gamma = div A.

System I:  dA^i/dt = -E^i - d_i psi
           dE^i/dt = -d_j d_j A^i + d_j d_i A^j
           dpsi/dt = -d_i A^i
System II: dA^i/dt = -E^i - d_i psi,     dE^i/dt = d_i Gamma - d_j d_j A^i
           dpsi/dt = -Gamma,             dGamma/dt = -d_i d_i psi
Constraints: C = d_i E^i (both systems), G = Gamma - d_i A^i (System II).
``Maxwell.InitialData::System_to_use`` selects the system.
"""
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import grid as gri
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from . import InitialData  # noqa: F401  (registers System_to_use)


def VacuumMaxwellRHSs():
    global ArhsU, ErhsU, C, psi_rhs
    system = par.parval_from_str("Maxwell.InitialData::System_to_use")
    rfm.reference_metric()
    EU = ixp.register_gridfunctions_for_single_rank1("EVOL", "EU")
    ixp.register_gridfunctions_for_single_rank1("EVOL", "AU")
    gri.register_gridfunctions("EVOL", ["psi"])
    psi_dD = ixp.declarerank1("psi_dD")
    AU_dD = ixp.declarerank2("AU_dD", "nosym")       # AU_dD[i][k] = d_k A^i
    AU_dDD = ixp.declarerank3("AU_dDD", "sym12")     # AU_dDD[i][k][m] = d_k d_m A^i
    EU_dD = ixp.declarerank2("EU_dD", "nosym")
    gri.register_gridfunctions("AUX", "DivE")
    C = sum(EU_dD[i][i] for i in range(3))
    ArhsU = [-EU[i] - psi_dD[i] for i in range(3)]
    if system == "System_I":
        psi_rhs = -sum(AU_dD[i][i] for i in range(3))
        ErhsU = [sum(-AU_dDD[i][j][j] + AU_dDD[j][j][i] for j in range(3)) for i in range(3)]
    elif system == "System_II":
        global Gamma_rhs, G
        Gamma = gri.register_gridfunctions("EVOL", ["Gamma"])
        Gamma_dD = ixp.declarerank1("Gamma_dD")
        psi_dDD = ixp.declarerank2("psi_dDD", "sym01")
        psi_rhs = -Gamma
        gri.register_gridfunctions("AUX", ["G"])
        G = Gamma - sum(AU_dD[i][i] for i in range(3))
        Gamma_rhs = -sum(psi_dDD[i][i] for i in range(3))
        ErhsU = [Gamma_dD[i] - sum(AU_dDD[i][j][j] for j in range(3)) for i in range(3)]
    else:
        raise ValueError("System_to_use must be System_I or System_II, not " + str(system))
