"""Scalar wave equation in a curvilinear reference metric.

v_t = c^2 (ghat^ij u_,ij - ghat^ij Gammahat^k_ij u_,k)
(ScalarWave/ScalarWaveCurvilinear_RHSs.py:43-98)
"""
from .. import NRPy_param_funcs as par
from .. import grid as gri
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from .CommonParams import wavespeed


def ScalarWaveCurvilinear_RHSs():
    global uu_rhs, vv_rhs
    DIM = par.parval_from_str("grid::DIM")
    rfm.reference_metric()
    contractedGammahatU = ixp.zerorank1()
    for k in range(DIM):
        for i in range(DIM):
            for j in range(DIM):
                contractedGammahatU[k] += rfm.ghatUU[i][j] * rfm.GammahatUDD[k][i][j]
    if any(gf.name == "uu" for gf in gri.glb_gridfcs_list):
        import sympy as sp
        uu, vv = sp.symbols("uu vv", real=True)
    else:
        uu, vv = gri.register_gridfunctions("EVOL", ["uu", "vv"])
    uu_dD = ixp.declarerank1("uu_dD")
    uu_dDD = ixp.declarerank2("uu_dDD", "sym01")
    uu_rhs = vv
    s = 0
    for i in range(DIM):
        s -= contractedGammahatU[i] * uu_dD[i]
        for j in range(DIM):
            s += rfm.ghatUU[i][j] * uu_dDD[i][j]
    vv_rhs = s * wavespeed * wavespeed
