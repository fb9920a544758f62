"""Scalar-wave initial data (ScalarWave/InitialData.py:28-102).

PlaneWave:        u = 2 + sin(k.x/|k| - c t),  v = du/dt
SphericalGaussian: u = 1 + [(r - ct) e^{-(r-ct)^2/(2 s^2)} + (r + ct) e^{-(r+ct)^2/(2 s^2)}]/r
"""
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import reference_metric as rfm
from .CommonParams import wavespeed

thismodule = "ScalarWave.InitialData"


def SphericalGaussian(CoordSystem="Cartesian", default_time=0, default_sigma=3):
    global uu_ID, vv_ID
    par.set_parval_from_str("reference_metric::CoordSystem", CoordSystem)
    rfm.reference_metric()
    time = par.Cparameters("REAL", thismodule, "time", default_time)
    sigma = par.Cparameters("REAL", thismodule, "sigma", default_sigma)
    r = sp.sqrt(sum(rfm.xx_to_Cart[i] ** 2 for i in range(3)))
    rm, rp = r - wavespeed * time, r + wavespeed * time
    uu_ID = rm / r * sp.exp(-rm ** 2 / (2 * sigma ** 2)) + rp / r * sp.exp(-rp ** 2 / (2 * sigma ** 2)) + 1
    vv_ID = sp.diff(uu_ID, time)


def PlaneWave(CoordSystem="Cartesian", default_time=0, default_k0=1, default_k1=1, default_k2=1):
    global uu_ID, vv_ID
    par.set_parval_from_str("reference_metric::CoordSystem", CoordSystem)
    rfm.reference_metric()
    time = par.Cparameters("REAL", thismodule, "time", default_time)
    kk = par.Cparameters("REAL", thismodule, ["kk0", "kk1", "kk2"], [default_k0, default_k1, default_k2])
    knorm = sp.sqrt(kk[0] ** 2 + kk[1] ** 2 + kk[2] ** 2)
    dot = sum(kk[i] * rfm.xx_to_Cart[i] for i in range(3)) / knorm
    uu_ID = sp.sin(dot - wavespeed * time) + 2
    vv_ID = sp.diff(uu_ID, time)


def InitialData(Type="PlaneWave", CoordSystem="Cartesian", default_time=0,
                default_k0=1, default_k1=1, default_k2=1, default_sigma=3):
    if Type == "PlaneWave":
        PlaneWave(CoordSystem, default_time, default_k0, default_k1, default_k2)
    elif Type == "SphericalGaussian":
        SphericalGaussian(CoordSystem, default_time, default_sigma)
    else:
        raise ValueError("ScalarWave initial data Type=" + str(Type) + " not supported")
