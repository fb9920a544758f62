"""Toroidal dipole initial data for vacuum Maxwell (Maxwell/InitialData.py:25-113;
gr-qc/0201051 eq. 16).

In spherical polar components, with u = t + r, v = t - r,
  A_phihat = amp sin(theta) [ (e^{-lam v^2} - e^{-lam u^2}) / r^2
                               - 2 lam (v e^{-lam v^2} + u e^{-lam u^2}) / r ],
and A_phi = A_phihat / (r sin(theta)); the data are built in Cartesian
coordinates, then transformed to the basis of reference_metric::CoordSystem.
E^i = -dA^i/dt, psi = 0 (and Gamma = 0 for System II).
"""
import sympy as sp
from .. import NRPy_param_funcs as par
from .. import indexedexp as ixp
from .. import reference_metric as rfm
from .CommonParams import amp, lam, time

thismodule = "Maxwell.InitialData"
par.initialize_param(par.glb_param("char", thismodule, "System_to_use", "System_I"))


def Toroidal():
    global AidU, EidU, psi_ID
    system = par.parval_from_str(thismodule + "::System_to_use")
    dst = par.parval_from_str("reference_metric::CoordSystem")
    par.set_parval_from_str("reference_metric::CoordSystem", "Cartesian")
    rfm.reference_metric()
    x, y, z = rfm.xx_to_Cart[0], rfm.xx_to_Cart[1], rfm.xx_to_Cart[2]
    r = sp.sqrt(x * x + y * y + z * z)
    sin_theta = z / r
    u, v = time + r, time - r
    eu, ev = sp.exp(-lam * u ** 2), sp.exp(-lam * v ** 2)
    A_phihat = amp * sin_theta * ((ev - eu) / r ** 2 - 2 * lam * (v * ev + u * eu) / r)
    A_Sph_D = [sp.sympify(0), sp.sympify(0), A_phihat / (r * sin_theta)]
    # d xSph / d xCart, inverted: d xCart / d xSph
    JsphCart = [[sp.diff(rfm.xxSph[i], rfm.xx_to_Cart[j]) for j in range(3)] for i in range(3)]
    JcartSph, _ = ixp.generic_matrix_inverter3x3(JsphCart)
    A_Cart = [sp.simplify(sum(JcartSph[i][j] * A_Sph_D[j] for j in range(3))) for i in range(3)]
    cart_xx = ixp.declarerank1("cart_xx")
    A_Cart = [e.subs({rfm.xx[k]: cart_xx[k] for k in range(3)}, simultaneous=True) for e in A_Cart]
    par.set_parval_from_str("reference_metric::CoordSystem", dst)
    rfm.reference_metric()
    A_Cart = [e.subs({cart_xx[k]: rfm.xx_to_Cart[k] for k in range(3)}, simultaneous=True) for e in A_Cart]
    _, Jinv = rfm.compute_Jacobian_and_inverseJacobian_tofrom_Cartesian()
    AidU = rfm.basis_transform_vectorU_from_Cartesian_to_rfmbasis(Jinv, A_Cart)
    EidU = [-sp.diff(AidU[j], time) for j in range(3)]
    psi_ID = sp.sympify(0)
    if system == "System_II":
        global Gamma_ID
        Gamma_ID = sp.sympify(0)
    elif system != "System_I":
        raise ValueError("System_to_use must be System_I or System_II, not " + str(system))


def InitialData(Type="ToroidalDipole"):
    if Type != "ToroidalDipole":
        raise ValueError("vacuum Maxwell initial data Type=" + str(Type) + " not supported")
    Toroidal()
