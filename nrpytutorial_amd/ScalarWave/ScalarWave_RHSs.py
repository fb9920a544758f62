"""Scalar wave equation as a first-order-in-time system, Cartesian coordinates.

u_t = v,  v_t = c^2 sum_i u_,ii   (ScalarWave/ScalarWave_RHSs.py:46-71)
"""
from .. import NRPy_param_funcs as par
from .. import grid as gri
from .. import indexedexp as ixp
from .CommonParams import wavespeed


def ScalarWave_RHSs():
    global uu_rhs, vv_rhs
    DIM = par.parval_from_str("grid::DIM")
    if any(gf.name == "uu" for gf in gri.glb_gridfcs_list):
        import sympy as sp
        uu, vv = sp.symbols("uu vv", real=True)
    else:
        uu, vv = gri.register_gridfunctions("EVOL", ["uu", "vv"])
    uu_dDD = ixp.declarerank2("uu_dDD", "sym01")
    uu_rhs = vv
    vv_rhs = 0
    for i in range(DIM):
        vv_rhs += wavespeed * wavespeed * uu_dDD[i][i]
