"""Host-side driver for a compiled system bundle, through its C ABI (ctypes).

This plays the role of the reference's Start-to-Finish ``main()``
(Tutorial-Start_to_Finish-BSSNCurvilinear-Two_BHs_Collide.ipynb:975-1187):
grid setup with the EigenCoord grid for boundary maps and the chosen
coordinate grid for evolution, precomputed reference-metric tables, initial
data + boundary conditions at t=0, the RK4 time loop and the Hamiltonian
constraint diagnostic. Device arrays are torch tensors (HBM allocation and
streams only); every computation is a call into the bundle's HIP kernels.
There is no CPU fallback: a missing library or device raises.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
GENDIR = os.path.join(HERE, "csrc", "generated")

NRPY_ERRORS = {-1: "NULL pointer", -2: "inconsistent grid sizes", -3: "kernel launch / HIP runtime failure",
               -4: "invalid argument", -5: "allocation failed", -6: "boundary map: Cartesian disagreement or bad parity",
               -7: "request exceeds a fixed capacity"}


class NrpyError(RuntimeError):
    pass


def check(rc, what):
    if rc != 0:
        raise NrpyError("%s failed with status %d (%s)" % (what, rc, NRPY_ERRORS.get(rc, "unknown")))


class OuterBC(ctypes.Structure):
    _fields_ = [("i0", ctypes.c_short), ("i1", ctypes.c_short), ("i2", ctypes.c_short),
                ("FACEi0", ctypes.c_int8), ("FACEi1", ctypes.c_int8), ("FACEi2", ctypes.c_int8)]


class InnerBC(ctypes.Structure):
    _fields_ = [("d0", ctypes.c_short), ("d1", ctypes.c_short), ("d2", ctypes.c_short),
                ("s0", ctypes.c_short), ("s1", ctypes.c_short), ("s2", ctypes.c_short),
                ("parity", ctypes.c_int8 * 10)]


_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_DPP = ctypes.POINTER(ctypes.c_void_p)

# exported symbol -> (restype, argtypes); params/rfm/bc are opaque pointers here
PROTOTYPES = {
    "set_Nxx_dxx_invdx_params__and__xx": (_I, [_I, ctypes.POINTER(_I), _VP, _DPP]),
    "xx_to_Cart": (None, [_VP, _DPP, _I, _I, _I, ctypes.POINTER(_D)]),
    "find_timestep": (_D, [_VP, _DPP, _D]),
    "set_up__bc_gz_map_and_parity_condns": (_I, [_VP, _DPP, _VP, _VP]),
    "bcstruct_create": (_I, [_VP, _DPP, ctypes.POINTER(_VP)]),
    "bcstruct_create_host": (_I, [_VP, _DPP, ctypes.POINTER(_VP)]),
    "bcstruct_free": (None, [_VP]),
    "bcstruct_get_counts": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "bcstruct_get_layer": (_I, [_VP, _I, _VP, _VP]),
    "bcstruct_get_composed": (_I, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "apply_bcs_curvilinear": (_I, [_VP, _VP, _I, _VP, _VP, _VP]),
    "apply_bcs_curvilinear_composed": (_I, [_VP, _VP, _I, _VP, _VP, _VP]),
    "gridfunction_sumsq": (_I, [_VP, _VP, _I, _VP, _VP, _VP]),
    "gridfunction_sumsq_noalloc": (_I, [_VP, _VP, _I, _VP, _VP]),
    "rfm_struct_create": (_I, [_VP, _DPP, ctypes.POINTER(_VP)]),
    "rfm_struct_free": (None, [_VP]),
    "rhs_eval_rk4_substep": (_I, [_I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _D, _VP]),
    "MoL_step_forward_in_time": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _D, _VP]),
    "MoL_graph_create": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _D, ctypes.POINTER(_VP)]),
    "MoL_graph_launch": (_I, [_VP, _I, _VP]),
    "MoL_graph_free": (None, [_VP]),
    "initial_data": (_I, [_VP, _DPP, _VP, _VP]),
    "nrpy_rk4_update_all": (_I, [_I, ctypes.c_size_t, _VP, _VP, _VP, _D, _VP]),
}
# Sommerfeld radiation condition (bundles built with it): apply_bcs_sommerfeld(params, xx, bc,
# NUM_GFS, gfs_parity, gfs, rhs_gfs, stream) and the RK4 step that uses it.
SOMMERFELD_PROTOTYPES = {
    "apply_bcs_sommerfeld": (_I, [_VP, _DPP, _VP, _I, _VP, _VP, _VP, _VP]),
    "MoL_step_forward_in_time_sommerfeld": (_I, [_VP, _VP, _DPP, _VP, _VP, _VP, _VP, _VP, _D, _VP]),
    "MoL_step_forward_in_time_sommerfeld_unfused": (_I, [_VP, _VP, _DPP, _VP, _VP, _VP, _VP, _VP, _D, _VP]),
    "MoL_sommerfeld_graph_create": (_I, [_VP, _VP, _DPP, _VP, _VP, _VP, _VP, _VP, _D, ctypes.POINTER(_VP)]),
}
KERNEL_PROTOTYPES = {
    # reference-style kernels: (params..., stream)
    "scalarwave": {"rhs_eval": [_VP, _VP, _VP, _VP]},
    "scalarwave_curvilinear": {"rhs_eval": [_VP, _VP, _VP, _VP, _VP]},
    "bssn": {"rhs_eval": [_VP, _VP, _VP, _VP, _VP, _VP], "Ricci_eval": [_VP, _VP, _VP, _VP, _VP],
             "rhs_eval_vjp": [_VP] * 9, "Ricci_eval_vjp": [_VP] * 9,
             "Hamiltonian_constraint": [_VP, _VP, _VP, _VP, _VP],
             "momentum_constraint": [_VP, _VP, _VP, _VP, _VP],
             "enforce_detgammabar_constraint": [_VP, _VP, _VP, _VP]},
    "maxwell": {"rhs_eval": [_VP, _VP, _VP, _VP], "Maxwell_constraints": [_VP, _VP, _VP, _VP]},
    "maxwell_curvilinear": {"rhs_eval": [_VP, _VP, _VP, _VP, _VP], "Maxwell_constraints": [_VP, _VP, _VP, _VP, _VP]},
}


def load_meta(name):
    with open(os.path.join(GENDIR, name + ".json")) as f:
        return json.load(f)


def make_paramstruct_type(members):
    fields = [(n, ctypes.c_int if t == "int" else ctypes.c_double) for n, t, _ in members]
    return type("paramstruct", (ctypes.Structure,), {"_fields_": fields})


class Bundle:
    """A compiled system bundle (``lib/libnrpy_<name>.so``)."""

    def __init__(self, name):
        self.name = name
        self.meta = load_meta(name)
        path = os.path.join(LIBDIR, "libnrpy_%s.so" % name)
        if not os.path.exists(path):
            raise NrpyError("HIP library %s is missing: run __graft_entry__.build() first" % path)
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.Params = make_paramstruct_type(self.meta["paramstruct"])
        self.NGHOSTS = self.meta["NGHOSTS"]
        self.evol = self.meta["gfs"]["evol"]
        self.aux = self.meta["gfs"]["aux"]
        self.auxevol = self.meta["gfs"]["auxevol"]
        self.evol_parity = (ctypes.c_int8 * max(1, len(self.evol)))(*self.meta["evol_parity"])
        for sym, (res, args) in PROTOTYPES.items():
            f = getattr(self.lib, sym)
            f.restype = res
            f.argtypes = args
        if "sommerfeld" in self.meta:
            for sym, (res, args) in SOMMERFELD_PROTOTYPES.items():
                f = getattr(self.lib, sym)
                f.restype = res
                f.argtypes = args
        family = self.meta.get("family", name.split("_")[0])
        for sym, args in KERNEL_PROTOTYPES.get(family, {}).items():
            f = getattr(self.lib, sym)
            f.restype = _I
            f.argtypes = args

    def new_params(self, **overrides):
        p = self.Params()
        for n, t, d in self.meta["paramstruct"]:
            setattr(p, n, int(d) if t == "int" else float(d))
        for k, v in overrides.items():
            if not hasattr(p, k):
                raise KeyError("paramstruct has no member " + k)
            setattr(p, k, v)
        return p

    def gf(self, name):
        return self.evol.index(name)


def _xx_ptrs(xx):
    arr = (ctypes.c_void_p * 3)()
    for d in range(3):
        arr[d] = xx[d].ctypes.data
    return arr


def host_grid(bundle, Nxx, eigen, **params):
    """(paramstruct, [xx0, xx1, xx2] numpy arrays) from set_Nxx_dxx_invdx_params__and__xx."""
    p = bundle.new_params(**params)
    G = bundle.NGHOSTS
    xx = [np.zeros(int(Nxx[d]) + 2 * G) for d in range(3)]
    n = (ctypes.c_int * 3)(*[int(v) for v in Nxx])
    check(bundle.lib.set_Nxx_dxx_invdx_params__and__xx(1 if eigen else 0, n, ctypes.byref(p), _xx_ptrs(xx)),
          "set_Nxx_dxx_invdx_params__and__xx")
    return p, xx


class HostBC:
    """Boundary lists built on the host only (no device memory)."""

    def __init__(self, bundle, Nxx, **params):
        self.bundle = bundle
        pe, xxe = host_grid(bundle, Nxx, True, **params)
        self.params = pe
        h = ctypes.c_void_p()
        check(bundle.lib.bcstruct_create_host(ctypes.byref(pe), _xx_ptrs(xxe), ctypes.byref(h)), "bcstruct_create_host")
        self.h = h

    def counts(self):
        G = self.bundle.NGHOSTS
        ob, ib = (ctypes.c_int * G)(), (ctypes.c_int * G)()
        ng, nt = ctypes.c_int(), ctypes.c_int()
        check(self.bundle.lib.bcstruct_get_counts(self.h, ob, ib, ctypes.byref(ng), ctypes.byref(nt)), "counts")
        return list(ob), list(ib), ng.value, nt.value

    def layer(self, gz):
        ob, ib, _, _ = self.counts()
        o = (OuterBC * max(1, ob[gz]))()
        i = (InnerBC * max(1, ib[gz]))()
        check(self.bundle.lib.bcstruct_get_layer(self.h, gz, o, i), "bcstruct_get_layer")
        outer = [(x.i0, x.i1, x.i2, x.FACEi0, x.FACEi1, x.FACEi2) for x in o[:ob[gz]]]
        inner = [(x.d0, x.d1, x.d2, x.s0, x.s1, x.s2) + tuple(x.parity) for x in i[:ib[gz]]]
        return outer, inner

    def composed(self):
        _, _, ng, nt = self.counts()
        gdest = np.zeros(ng, np.int32)
        rowptr = np.zeros(ng + 1, np.int32)
        src = np.zeros(max(1, nt), np.int32)
        coeff = np.zeros(max(1, nt))
        neg = np.zeros(max(1, nt), np.int32)
        check(self.bundle.lib.bcstruct_get_composed(self.h, gdest.ctypes.data, rowptr.ctypes.data, src.ctypes.data,
                                                    coeff.ctypes.data, neg.ctypes.data), "bcstruct_get_composed")
        return gdest, rowptr, src[:nt], coeff[:nt], neg[:nt]

    def __del__(self):
        try:
            self.bundle.lib.bcstruct_free(self.h)
        except Exception:
            pass


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise NrpyError("no GPU visible: the HIP kernels need a device")
    return torch


def stream_handle(torch, stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Grid:
    """Device-side grid: paramstruct, coordinates, boundary struct, rfm tables."""

    def __init__(self, bundle, Nxx, **params):
        torch = _torch()
        self.torch = torch
        self.bundle = bundle
        self.Nxx = tuple(int(v) for v in Nxx)
        pe, xxe = host_grid(bundle, Nxx, True, **params)
        bc = ctypes.c_void_p()
        check(bundle.lib.bcstruct_create(ctypes.byref(pe), _xx_ptrs(xxe), ctypes.byref(bc)), "bcstruct_create")
        self.bc = bc
        self.params, self.xx = host_grid(bundle, Nxx, False, **params)
        rfmp = ctypes.c_void_p()
        check(bundle.lib.rfm_struct_create(ctypes.byref(self.params), _xx_ptrs(self.xx), ctypes.byref(rfmp)),
              "rfm_struct_create")
        self.rfm = rfmp
        self.xx_dev = [torch.from_numpy(x).cuda() for x in self.xx]
        self.xx_dev_ptrs = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in self.xx_dev])
        G = bundle.NGHOSTS
        self.shape = (self.Nxx[2] + 2 * G, self.Nxx[1] + 2 * G, self.Nxx[0] + 2 * G)
        self.Ntot = self.shape[0] * self.shape[1] * self.shape[2]

    def pp(self):
        return ctypes.byref(self.params)

    def find_timestep(self, CFL):
        return self.bundle.lib.find_timestep(self.pp(), _xx_ptrs(self.xx), float(CFL))

    def empty(self, ngf):
        return self.torch.zeros((max(1, ngf),) + self.shape, dtype=self.torch.float64, device="cuda")

    def interior(self, t):
        G = self.bundle.NGHOSTS
        return t[..., G:G + self.Nxx[2], G:G + self.Nxx[1], G:G + self.Nxx[0]]

    def __del__(self):
        try:
            self.bundle.lib.bcstruct_free(self.bc)
            self.bundle.lib.rfm_struct_free(self.rfm)
        except Exception:
            pass


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Evolution:
    """RK4 method-of-lines evolution on one grid (arrays of RK_Allocate_Memory.h,
    MoLtimestepping/C_Code_Generation.py:68-105)."""

    def __init__(self, grid):
        self.g = grid
        b = grid.bundle
        self.lib = b.lib
        n = len(b.evol)
        self.y_n = grid.empty(n)
        self.k_odd = grid.empty(n)
        self.k_even = grid.empty(n)
        self.rt = grid.empty(n)
        self.auxevol = grid.empty(len(b.auxevol))
        self.aux = grid.empty(len(b.aux))
        self.workspace = grid.torch.zeros(1024 + 1, dtype=grid.torch.float64, device="cuda")
        self.graph = None
        self.graph_dt = None
        self.som_graph = None
        self.som_graph_dt = None

    def s(self):
        return stream_handle(self.g.torch)

    def initial_data(self, apply_bcs=True):
        check(self.lib.initial_data(self.g.pp(), self.g.xx_dev_ptrs, ptr(self.y_n), self.s()), "initial_data")
        if apply_bcs:
            self.apply_bcs(self.y_n)
            self.post_bcs(self.y_n)

    def apply_bcs(self, gfs, composed=False):
        f = self.lib.apply_bcs_curvilinear_composed if composed else self.lib.apply_bcs_curvilinear
        check(f(self.g.pp(), self.g.bc, len(self.g.bundle.evol), self.g.bundle.evol_parity, ptr(gfs), self.s()),
              "apply_bcs_curvilinear")

    def post_bcs(self, gfs):
        if self.g.bundle.meta.get("has_detg"):
            check(self.lib.enforce_detgammabar_constraint(self.g.rfm, self.g.pp(), ptr(gfs), self.s()),
                  "enforce_detgammabar_constraint")

    def step(self, dt, nsteps=1, use_graph=True):
        """Advance nsteps RK4 steps of size dt."""
        if use_graph:
            if self.graph is None or self.graph_dt != dt:
                self.free_graph()
                h = ctypes.c_void_p()
                check(self.lib.MoL_graph_create(self.g.rfm, self.g.pp(), self.g.bc, ptr(self.y_n), ptr(self.k_odd),
                                                ptr(self.k_even), ptr(self.rt), ptr(self.auxevol), float(dt),
                                                ctypes.byref(h)), "MoL_graph_create")
                self.graph, self.graph_dt = h, dt
            check(self.lib.MoL_graph_launch(self.graph, int(nsteps), self.s()), "MoL_graph_launch")
        else:
            for _ in range(nsteps):
                check(self.lib.MoL_step_forward_in_time(self.g.rfm, self.g.pp(), self.g.bc, ptr(self.y_n),
                                                        ptr(self.k_odd), ptr(self.k_even), ptr(self.rt),
                                                        ptr(self.auxevol), float(dt), self.s()),
                      "MoL_step_forward_in_time")

    def apply_bcs_sommerfeld(self, gfs, rhs_gfs):
        """Sommerfeld radiation condition on rhs_gfs at every ghost point (bundles built with it)."""
        check(self.lib.apply_bcs_sommerfeld(self.g.pp(), self.g.xx_dev_ptrs, self.g.bc, len(self.g.bundle.evol),
                                            self.g.bundle.evol_parity, ptr(gfs), ptr(rhs_gfs), self.s()),
              "apply_bcs_sommerfeld")

    def step_sommerfeld(self, dt, nsteps=1, use_graph=True, unfused=False):
        """nsteps RK4 steps with the Sommerfeld condition on the RHS and the update at every point,
        replayed from a captured hipGraph (or launched eagerly). unfused: launch the step in the
        reference's order of operations (MoL_step_forward_in_time_sommerfeld_unfused) eagerly;
        the default step is fused where the grid allows it and gives the same values."""
        if unfused:
            for _ in range(nsteps):
                check(self.lib.MoL_step_forward_in_time_sommerfeld_unfused(
                    self.g.pp(), self.g.rfm, self.g.xx_dev_ptrs, self.g.bc, ptr(self.y_n), ptr(self.k_odd),
                    ptr(self.k_even), ptr(self.rt), float(dt), self.s()), "MoL_step_forward_in_time_sommerfeld_unfused")
            return
        if use_graph:
            if self.som_graph is None or self.som_graph_dt != dt:
                self.free_graph()
                h = ctypes.c_void_p()
                check(self.lib.MoL_sommerfeld_graph_create(self.g.pp(), self.g.rfm, self.g.xx_dev_ptrs, self.g.bc,
                                                           ptr(self.y_n), ptr(self.k_odd), ptr(self.k_even),
                                                           ptr(self.rt), float(dt), ctypes.byref(h)),
                      "MoL_sommerfeld_graph_create")
                self.som_graph, self.som_graph_dt = h, dt
            check(self.lib.MoL_graph_launch(self.som_graph, int(nsteps), self.s()), "MoL_graph_launch")
            return
        for _ in range(nsteps):
            check(self.lib.MoL_step_forward_in_time_sommerfeld(self.g.pp(), self.g.rfm, self.g.xx_dev_ptrs, self.g.bc,
                                                               ptr(self.y_n), ptr(self.k_odd), ptr(self.k_even),
                                                               ptr(self.rt), float(dt), self.s()),
                  "MoL_step_forward_in_time_sommerfeld")

    def momentum(self, gfs=None):
        """The momentum constraint M^i (rescaled, three gridfunctions) at interior points."""
        gfs = self.y_n if gfs is None else gfs
        check(self.lib.momentum_constraint(self.g.rfm, self.g.pp(), ptr(gfs), ptr(self.aux), self.s()),
              "momentum_constraint")
        a = self.g.bundle.aux
        return self.aux[a.index("MU0"):a.index("MU2") + 1]

    def hamiltonian(self, gfs=None):
        gfs = self.y_n if gfs is None else gfs
        check(self.lib.Hamiltonian_constraint(self.g.rfm, self.g.pp(), ptr(gfs), ptr(self.aux), self.s()),
              "Hamiltonian_constraint")
        return self.aux[self.g.bundle.aux.index("H")]

    def sumsq(self, gfs, gf):
        res = self.workspace[1024:]
        check(self.lib.gridfunction_sumsq(self.g.pp(), ptr(gfs), int(gf), ptr(self.workspace), ptr(res), self.s()),
              "gridfunction_sumsq")
        return res

    def free_graph(self):
        if self.graph is not None:
            self.lib.MoL_graph_free(self.graph)
            self.graph = None
        if self.som_graph is not None:
            self.lib.MoL_graph_free(self.som_graph)
            self.som_graph = None

    def __del__(self):
        try:
            self.free_graph()
        except Exception:
            pass
