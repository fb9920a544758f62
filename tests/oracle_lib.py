"""ctypes access to the C oracle (oracle/generated/liboracle_<bundle>.so).

Test infrastructure: used by the parity tests, __graft_entry__.smoke() and
bench.py's cpu_baseline leg only.
"""
import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "oracle", "generated")

_VP, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_DPP = ctypes.POINTER(ctypes.c_void_p)


class Oracle:
    def __init__(self, bundle, fast=False, variant=None):
        """variant: None, or "asan" for the sanitizer build (oracle/Makefile target asan)."""
        self.meta = json.load(open(os.path.join(GEN, bundle + ".json")))
        kind = variant + "_" if variant else "fast_" if fast else ""
        path = os.path.join(GEN, "liboracle_%s%s.so" % (kind, bundle))
        if not os.path.exists(path):
            raise RuntimeError("oracle library %s missing: build first" % path)
        self.lib = ctypes.CDLL(path)
        fields = [(m[0], ctypes.c_int if m[1] == "int" else ctypes.c_double) for m in self.meta["paramstruct"]]
        self.Params = type("oparams", (ctypes.Structure,), {"_fields_": fields})
        L = self.lib
        L.oracle_set_Nxx_dxx_invdx_params__and__xx.argtypes = [_I, ctypes.POINTER(_I), _VP, _DPP]
        L.oracle_bcstruct_create.argtypes = [_VP, _DPP]
        L.oracle_bcstruct_create.restype = _VP
        L.oracle_bcstruct_free.argtypes = [_VP]
        L.oracle_bcstruct_counts.argtypes = [_VP, _VP, _VP]
        L.oracle_bcstruct_layer.argtypes = [_VP, _I, _VP, _VP]
        L.oracle_apply_bcs.argtypes = [_VP, _VP, _I, _VP, _VP]
        L.oracle_rhs_eval.argtypes = [_VP, _DPP, _VP, _VP, _VP]
        L.oracle_Ricci_eval.argtypes = [_VP, _DPP, _VP, _VP]
        L.oracle_enforce_detgammabar_constraint.argtypes = [_VP, _DPP, _VP]
        L.oracle_initial_data.argtypes = [_VP, _DPP, _VP]
        L.oracle_MoL_step.argtypes = [_VP, _DPP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _D]
        L.oracle_find_timestep.argtypes = [_VP, _DPP, _D]
        L.oracle_find_timestep.restype = _D
        L.oracle_xx_to_Cart.argtypes = [_VP, _DPP, _I, _I, _I, _VP]
        if hasattr(L, "oracle_Maxwell_constraints"):
            L.oracle_Maxwell_constraints.argtypes = [_VP, _DPP, _VP, _VP]
        if hasattr(L, "oracle_Hamiltonian_constraint"):
            L.oracle_Hamiltonian_constraint.argtypes = [_VP, _DPP, _VP, _VP]
        if hasattr(L, "oracle_momentum_constraint"):
            L.oracle_momentum_constraint.argtypes = [_VP, _DPP, _VP, _VP]
        # exact-derivative twins: derivative values in the order of meta["exact_derivs"][<function>]
        if hasattr(L, "oracle_rhs_eval_exact"):
            L.oracle_Ricci_eval_exact.argtypes = [_VP, _DPP, _VP, _VP, _VP]
            L.oracle_rhs_eval_exact.argtypes = [_VP, _DPP, _VP, _VP, _VP, _VP]
            L.oracle_Hamiltonian_constraint_exact.argtypes = [_VP, _DPP, _VP, _VP, _VP]
        self.sommerfeld = None
        if hasattr(L, "oracle_apply_bcs_sommerfeld"):
            L.oracle_apply_bcs_sommerfeld.argtypes = [_VP, _DPP, _VP, _I, _VP, _VP, _VP, _I, _VP, _VP, _VP]
            L.oracle_MoL_step_sommerfeld.argtypes = [_VP, _DPP, _VP, _VP, _VP, _VP, _VP, _VP, _D]
            out = (ctypes.c_double * 4)()
            L.oracle_sommerfeld_settings(out)
            self.sommerfeld = {"fd_order": int(out[0]), "at_inf": out[1], "radial_falloff_power": out[2],
                               "speed": out[3]}
        self.NGHOSTS = L.oracle_NGHOSTS()
        self.evol = self.meta["evol"]
        self.parity = (ctypes.c_int8 * len(self.evol))(*self.meta["evol_parity"])

    def params(self, **kw):
        p = self.Params()
        for n, t, d in self.meta["paramstruct"]:
            setattr(p, n, int(d) if t == "int" else float(d))
        for k, v in kw.items():
            if hasattr(p, k):
                setattr(p, k, v)
        return p


class OracleGrid:
    def __init__(self, oracle, Nxx, params_dict):
        self.o = oracle
        L = oracle.lib
        G = oracle.NGHOSTS
        self.Nxx = tuple(int(v) for v in Nxx)
        n = (ctypes.c_int * 3)(*self.Nxx)
        self.pe = oracle.params(**params_dict)
        self.xxe = [np.zeros(v + 2 * G) for v in self.Nxx]
        L.oracle_set_Nxx_dxx_invdx_params__and__xx(1, n, ctypes.byref(self.pe), self.ptrs(self.xxe))
        self.bc = L.oracle_bcstruct_create(ctypes.byref(self.pe), self.ptrs(self.xxe))
        if not self.bc:
            raise RuntimeError("oracle boundary map failed")
        self.p = oracle.params(**params_dict)
        self.xx = [np.zeros(v + 2 * G) for v in self.Nxx]
        L.oracle_set_Nxx_dxx_invdx_params__and__xx(0, n, ctypes.byref(self.p), self.ptrs(self.xx))
        self.shape = tuple(v + 2 * G for v in reversed(self.Nxx))
        self.ngf = len(oracle.evol)

    @staticmethod
    def ptrs(xx):
        a = (ctypes.c_void_p * 3)()
        for d in range(3):
            a[d] = xx[d].ctypes.data
        return a

    def zeros(self, ngf):
        return np.zeros((max(1, ngf),) + self.shape)

    def counts(self):
        G = self.o.NGHOSTS
        a, b = (ctypes.c_int * G)(), (ctypes.c_int * G)()
        self.o.lib.oracle_bcstruct_counts(self.bc, a, b)
        return list(a), list(b)

    def layer(self, gz):
        ob, ib = self.counts()
        o6 = np.zeros((max(1, ob[gz]), 6), np.int32)
        i16 = np.zeros((max(1, ib[gz]), 16), np.int32)
        self.o.lib.oracle_bcstruct_layer(self.bc, gz, o6.ctypes.data, i16.ctypes.data)
        return o6[:ob[gz]], i16[:ib[gz]]

    def apply_bcs(self, gfs, parity=None):
        par = self.o.parity if parity is None else parity
        self.o.lib.oracle_apply_bcs(ctypes.byref(self.p), self.bc, gfs.shape[0], par, gfs.ctypes.data)

    def rhs_eval(self, y, auxevol=None):
        out = np.zeros_like(y)
        aux = auxevol if auxevol is not None else np.zeros(1)
        self.o.lib.oracle_rhs_eval(ctypes.byref(self.p), self.ptrs(self.xx), aux.ctypes.data, y.ctypes.data,
                                   out.ctypes.data)
        return out

    def ricci(self, y):
        aux = self.zeros(6)
        self.o.lib.oracle_Ricci_eval(ctypes.byref(self.p), self.ptrs(self.xx), y.ctypes.data, aux.ctypes.data)
        return aux

    def hamiltonian(self, y):
        aux = self.zeros(1)
        self.o.lib.oracle_Hamiltonian_constraint(ctypes.byref(self.p), self.ptrs(self.xx), y.ctypes.data,
                                                 aux.ctypes.data)
        return aux[0]

    def exact_deriv_names(self, function):
        """Derivative symbols the *_exact twin of ``function`` reads, in deriv_gfs order."""
        return self.o.meta["exact_derivs"][function]

    def ricci_exact(self, y, derivs):
        aux = self.zeros(6)
        self.o.lib.oracle_Ricci_eval_exact(ctypes.byref(self.p), self.ptrs(self.xx), y.ctypes.data,
                                           np.ascontiguousarray(derivs).ctypes.data, aux.ctypes.data)
        return aux

    def rhs_eval_exact(self, y, auxevol, derivs):
        out = np.zeros_like(y)
        self.o.lib.oracle_rhs_eval_exact(ctypes.byref(self.p), self.ptrs(self.xx), auxevol.ctypes.data,
                                         y.ctypes.data, np.ascontiguousarray(derivs).ctypes.data, out.ctypes.data)
        return out

    def hamiltonian_exact(self, y, derivs):
        aux = self.zeros(1)
        self.o.lib.oracle_Hamiltonian_constraint_exact(ctypes.byref(self.p), self.ptrs(self.xx), y.ctypes.data,
                                                       np.ascontiguousarray(derivs).ctypes.data, aux.ctypes.data)
        return aux[0]

    def momentum(self, y):
        """M^0..M^2 (the AUX gridfunctions after H)."""
        aux = self.zeros(4)
        self.o.lib.oracle_momentum_constraint(ctypes.byref(self.p), self.ptrs(self.xx), y.ctypes.data,
                                              aux.ctypes.data)
        return aux[1:4]

    def maxwell_constraints(self, y):
        aux = self.zeros(2)
        self.o.lib.oracle_Maxwell_constraints(ctypes.byref(self.p), self.ptrs(self.xx), y.ctypes.data,
                                              aux.ctypes.data)
        return aux

    def enforce_detg(self, y):
        self.o.lib.oracle_enforce_detgammabar_constraint(ctypes.byref(self.p), self.ptrs(self.xx), y.ctypes.data)

    def initial_data(self, y):
        self.o.lib.oracle_initial_data(ctypes.byref(self.p), self.ptrs(self.xx), y.ctypes.data)

    def find_timestep(self, CFL):
        return self.o.lib.oracle_find_timestep(ctypes.byref(self.p), self.ptrs(self.xx), CFL)

    def mol_step(self, y_n, dt, nsteps=1):
        k1, k2, rt = np.zeros_like(y_n), np.zeros_like(y_n), np.zeros_like(y_n)
        aux = self.zeros(6)
        for _ in range(nsteps):
            self.o.lib.oracle_MoL_step(ctypes.byref(self.p), self.ptrs(self.xx), self.bc, self.o.parity,
                                       y_n.ctypes.data, k1.ctypes.data, k2.ctypes.data, rt.ctypes.data,
                                       aux.ctypes.data, dt)

    def apply_bcs_sommerfeld(self, y, rhs, fd_order=None, at_inf=None, power=None, speed=None):
        """Sommerfeld condition on rhs (in place) at every ghost point; parameters default to
        the bundle's Start-to-Finish settings, scalars apply to every gridfunction."""
        st = self.o.sommerfeld
        n = y.shape[0]
        arr = lambda v, d: np.full(n, float(d if v is None else v))
        a, pw, sp = arr(at_inf, st["at_inf"]), arr(power, st["radial_falloff_power"]), arr(speed, st["speed"])
        self.o.lib.oracle_apply_bcs_sommerfeld(ctypes.byref(self.p), self.ptrs(self.xx), self.bc, n,
                                               self.o.parity, y.ctypes.data, rhs.ctypes.data,
                                               int(fd_order or st["fd_order"]), a.ctypes.data, pw.ctypes.data,
                                               sp.ctypes.data)

    def mol_step_sommerfeld(self, y_n, dt, nsteps=1):
        k1, k2, rt = np.zeros_like(y_n), np.zeros_like(y_n), np.zeros_like(y_n)
        for _ in range(nsteps):
            self.o.lib.oracle_MoL_step_sommerfeld(ctypes.byref(self.p), self.ptrs(self.xx), self.bc, self.o.parity,
                                                  y_n.ctypes.data, k1.ctypes.data, k2.ctypes.data, rt.ctypes.data,
                                                  dt)

    def interior(self, a):
        G = self.o.NGHOSTS
        return a[..., G:G + self.Nxx[2], G:G + self.Nxx[1], G:G + self.Nxx[0]]

    def __del__(self):
        try:
            self.o.lib.oracle_bcstruct_free(self.bc)
        except Exception:
            pass
