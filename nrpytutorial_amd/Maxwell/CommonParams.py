"""C parameters shared by the vacuum Maxwell modules (Maxwell/CommonParams.py):
amplitude ``amp`` and width ``lam`` of the toroidal dipole data, ``time`` and
``wavespeed``."""
from .. import NRPy_param_funcs as par

thismodule = "Maxwell.CommonParams"
amp = par.Cparameters("REAL", thismodule, "amp", 1.0)
lam = par.Cparameters("REAL", thismodule, "lam", 1.0)
time = par.Cparameters("REAL", thismodule, "time", 0.0)
wavespeed = par.Cparameters("REAL", thismodule, "wavespeed", 1.0)
