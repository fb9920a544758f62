"""Parameters shared by the scalar-wave modules (ScalarWave/CommonParams.py:24)."""
from .. import NRPy_param_funcs as par

thismodule = "ScalarWave.CommonParams"
wavespeed = par.Cparameters("REAL", thismodule, "wavespeed", 1.0)
