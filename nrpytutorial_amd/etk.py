"""Host side of the Einstein Toolkit plugin form (bundle ``bssn_etk``).

The Baikal thorn's scheduled functions (BaikalETK/BaikalETK_C_drivers_codegen.py:
520-700) receive the grid and every gridfunction through CCTK_ARGUMENTS. Their
HIP counterparts in ``libnrpy_bssn_etk.so`` take the same information as plain C:

* ``nrpy_cctkGH``: ``cctk_lsh``, ``cctk_ash``, ``cctk_nghostzones`` and
  ``CCTK_DELTA_SPACE``;
* arrays of device pointers, one per gridfunction, in the gridfunction order of
  the bundle's metadata (EVOL: aDD00 ... vetU2; AUXEVOL: RbarDD00 ... RbarDD22;
  AUX: H, MU0, MU1, MU2), each array laid out as
  ``CCTK_GFINDEX3D(cctkGH, i, j, k) = i + ash0 (j + ash1 k)``.

This module mirrors those calls for Python callers (tests, benchmarks): each
gridfunction is a torch tensor of shape (ash2, ash1, ash0), or anything with a
``data_ptr()``. Status codes are raised as NrpyError, as in ``runtime``.
"""
import ctypes
import os

from .runtime import LIBDIR, NrpyError, check, load_meta, make_paramstruct_type

_VP, _I = ctypes.c_void_p, ctypes.c_int


class CctkGH(ctypes.Structure):
    """``nrpy_cctkGH``: what the kernels need from the toolkit's grid hierarchy."""
    _fields_ = [("cctk_lsh", ctypes.c_int * 3), ("cctk_ash", ctypes.c_int * 3),
                ("cctk_nghostzones", ctypes.c_int * 3), ("cctk_delta_space", ctypes.c_double * 3)]


def cctk_grid(lsh, nghostzones, delta_space, ash=None):
    """A CctkGH; ``ash`` defaults to ``lsh`` (no padding)."""
    g = CctkGH()
    for d in range(3):
        g.cctk_lsh[d] = lsh[d]
        g.cctk_ash[d] = (ash or lsh)[d]
        g.cctk_nghostzones[d] = nghostzones[d]
        g.cctk_delta_space[d] = delta_space[d]
    return g


_PROTOS = {
    "Baikal_BSSN_Ricci_FD_order_4": [_VP, _VP, _VP, _VP, _VP],
    "Baikal_BSSN_RHSs_enable_Tmunu_False_FD_order_4": [_VP, _VP, _VP, _VP, _VP, _VP],
    "Baikal_BSSN_constraints_FD_order_4": [_VP, _VP, _VP, _VP, _VP],
    "Baikal_enforce_detgammabar_constraint": [_VP, _VP, _VP, _VP],
    "Baikal_driver_pt1_BSSN_Ricci": [_VP, _VP, _I, _VP, _VP, _VP],
    "Baikal_driver_pt2_BSSN_RHSs": [_VP, _VP, _I, _VP, _VP, _VP, _VP],
    "Baikal_BSSN_constraints": [_VP, _VP, _I, _VP, _VP, _VP],
}


def _ptr(x):
    if x is None:
        return None
    return x if isinstance(x, int) else x.data_ptr()


class Baikal:
    """The BSSN kernels in the toolkit's plugin form."""

    def __init__(self, name="bssn_etk"):
        self.meta = load_meta(name)
        path = os.path.join(LIBDIR, "libnrpy_%s.so" % name)
        if not os.path.exists(path):
            raise NrpyError("HIP library %s is missing: run __graft_entry__.build() first" % path)
        self.lib = ctypes.CDLL(path)
        for sym, args in _PROTOS.items():
            f = getattr(self.lib, sym)
            f.restype = _I
            f.argtypes = args
        self.Params = make_paramstruct_type(self.meta["paramstruct"])
        self.evol = self.meta["gfs"]["evol"]
        self.auxevol = self.meta["gfs"]["auxevol"]
        self.aux = self.meta["gfs"]["aux"]
        self.FD_order = self.meta["FD_order"]
        self.NGHOSTS = self.meta["NGHOSTS"]

    def new_params(self, **overrides):
        p = self.Params()
        for n, t, d in self.meta["paramstruct"]:
            setattr(p, n, int(d) if t == "int" else float(d))
        for k, v in overrides.items():
            if not hasattr(p, k):
                raise KeyError("paramstruct has no member " + k)
            setattr(p, k, v)
        return p

    @staticmethod
    def pointers(arrays, count):
        """A C array of device pointers (one per gridfunction)."""
        if len(arrays) != count:
            raise NrpyError("expected %d gridfunction arrays, got %d" % (count, len(arrays)))
        return (_VP * count)(*[_ptr(a) for a in arrays])

    def ricci(self, gh, params, evol, rbar, stream=None, FD_order=None):
        """Baikal_driver_pt1_BSSN_Ricci: RbarDD at interior points."""
        check(self.lib.Baikal_driver_pt1_BSSN_Ricci(
            ctypes.byref(gh), ctypes.byref(params), self.FD_order if FD_order is None else FD_order,
            self.pointers(evol, len(self.evol)), self.pointers(rbar, len(self.auxevol)), stream),
            "Baikal_driver_pt1_BSSN_Ricci")

    def rhss(self, gh, params, evol, rbar, rhs, stream=None, FD_order=None):
        """Baikal_driver_pt2_BSSN_RHSs: every *_rhsGF at interior points."""
        check(self.lib.Baikal_driver_pt2_BSSN_RHSs(
            ctypes.byref(gh), ctypes.byref(params), self.FD_order if FD_order is None else FD_order,
            self.pointers(evol, len(self.evol)), self.pointers(rbar, len(self.auxevol)),
            self.pointers(rhs, len(self.evol)), stream), "Baikal_driver_pt2_BSSN_RHSs")

    def constraints(self, gh, params, evol, aux, stream=None, FD_order=None):
        """Baikal_BSSN_constraints: H, MU0, MU1, MU2 at interior points."""
        check(self.lib.Baikal_BSSN_constraints(
            ctypes.byref(gh), ctypes.byref(params), self.FD_order if FD_order is None else FD_order,
            self.pointers(evol, len(self.evol)), self.pointers(aux, len(self.aux)), stream),
            "Baikal_BSSN_constraints")

    def enforce_detgammabar(self, gh, params, evol, stream=None):
        """Baikal_enforce_detgammabar_constraint: rescale hDD at every point."""
        check(self.lib.Baikal_enforce_detgammabar_constraint(
            ctypes.byref(gh), ctypes.byref(params), self.pointers(evol, len(self.evol)), stream),
            "Baikal_enforce_detgammabar_constraint")
