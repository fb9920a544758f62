"""Many independent evolutions of one system, advanced together on one GPU.

Each instance has its own grid size, C parameters (e.g. black-hole masses and
positions), boundary maps, reference-metric tables, time step and state
arrays. ``BatchedEvolution`` hands all of them to ``nrpy_batch_create``; one
RK4 step of every instance is then one launch per kernel per substep
(``MoL_batch_step``), captured once into a hipGraph. ``autotune`` picks the
tile shape of the batched kernels by timing candidate shapes on the batch
itself. ``brill_lindquist_sweep`` draws the seeded parameter sweep of
head-on Brill-Lindquist collisions used to measure it.
"""
import ctypes
import math
import time

from .runtime import Grid, Evolution, NrpyError, check, ptr, stream_handle, _torch
from .workloads import uniforms

_VP = ctypes.c_void_p

# Tile shapes tried by autotune (threads per tile a multiple of 64, at most 256).
TILE_CANDIDATES = [(64, 4, 1), (64, 2, 2), (64, 2, 1), (64, 1, 2), (32, 8, 1), (32, 4, 2), (16, 16, 1), (16, 8, 2),
                   (128, 2, 1), (128, 1, 2)]


def brill_lindquist_sweep(n, seed=0, nr=(48, 96), nth=(8, 16)):
    """n head-on Brill-Lindquist setups: masses in [0.3,0.7], separation s in [0.6,2.0]
    (black holes at z = +-s/2), Nr and Ntheta even in the given ranges, Nphi = 2,
    RMAX = 7.5. Returns [(Nxx, params)]."""
    u = uniforms(seed, 5 * n)
    out = []
    for k in range(n):
        m1 = 0.3 + 0.4 * u[5 * k]
        m2 = 0.3 + 0.4 * u[5 * k + 1]
        sep = 0.6 + 1.4 * u[5 * k + 2]
        Nr = 2 * int(nr[0] // 2 + math.floor(u[5 * k + 3] * (nr[1] // 2 - nr[0] // 2 + 1)))
        Nth = 2 * int(nth[0] // 2 + math.floor(u[5 * k + 4] * (nth[1] // 2 - nth[0] // 2 + 1)))
        out.append(((Nr, Nth, 2), dict(BH1_mass=m1, BH2_mass=m2, BH1_posn_z=sep / 2, BH2_posn_z=-sep / 2,
                                        RMAX=7.5, eta=2.0)))
    return out


class BatchedEvolution:
    """Instances [(Nxx, params)] of one bundle, each with dt = find_timestep(CFL)."""

    def __init__(self, bundle, instances, CFL=0.5):
        self.torch = _torch()
        self.bundle = bundle
        self.lib = bundle.lib
        self.grids = [Grid(bundle, Nxx, **pars) for Nxx, pars in instances]
        self.evs = [Evolution(g) for g in self.grids]
        self.dts = [g.find_timestep(CFL) for g in self.grids]
        n = len(self.grids)
        self.n = n
        L = self.lib
        L.nrpy_batch_create.restype = ctypes.c_int
        L.nrpy_batch_create.argtypes = [ctypes.c_int, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(_VP)]
        L.nrpy_batch_free.argtypes = [_VP]
        L.nrpy_batch_prepare.restype = ctypes.c_int
        L.nrpy_batch_prepare.argtypes = [_VP]
        L.MoL_batch_step.restype = ctypes.c_int
        L.MoL_batch_step.argtypes = [_VP, _VP]
        L.MoL_batch_graph_create.restype = ctypes.c_int
        L.MoL_batch_graph_create.argtypes = [_VP, ctypes.POINTER(_VP)]
        self._params = (bundle.Params * n)(*[g.params for g in self.grids])
        self._rfms = (_VP * n)(*[g.rfm.value for g in self.grids])
        self._bcs = (_VP * n)(*[g.bc.value for g in self.grids])
        arrs = []
        for ev in self.evs:
            arrs += [ev.y_n.data_ptr(), ev.k_odd.data_ptr(), ev.k_even.data_ptr(), ev.rt.data_ptr(),
                     ev.auxevol.data_ptr(), ev.aux.data_ptr()]
        self._arrays = (_VP * (6 * n))(*arrs)
        self._dts = (ctypes.c_double * n)(*self.dts)
        h = _VP()
        check(L.nrpy_batch_create(n, self._params, self._rfms, self._bcs, self._arrays, self._dts, ctypes.byref(h)),
              "nrpy_batch_create")
        self.h = h
        self.graph = None

    def s(self):
        return stream_handle(self.torch)

    def initial_data(self):
        for ev in self.evs:
            ev.initial_data()

    def step_eager(self, nsteps=1):
        for _ in range(nsteps):
            check(self.lib.MoL_batch_step(self.h, self.s()), "MoL_batch_step")

    def step(self, nsteps=1):
        if self.graph is None:
            g = _VP()
            check(self.lib.MoL_batch_graph_create(self.h, ctypes.byref(g)), "MoL_batch_graph_create")
            self.graph = g
        check(self.lib.MoL_graph_launch(self.graph, int(nsteps), self.s()), "MoL_graph_launch")

    def free_graph(self):
        if self.graph is not None:
            self.lib.MoL_graph_free(self.graph)
            self.graph = None

    def hamiltonian_l2(self):
        """sqrt(mean H^2) over interior points, per instance."""
        out = []
        for ev in self.evs:
            H = ev.g.interior(ev.hamiltonian())
            out.append(float(self.torch.sqrt(self.torch.mean(H * H))))
        return out

    def set_tile(self, shape):
        """Tile shape of every batched kernel; the graph is rebuilt on the next step."""
        for name, kid in self.bundle.meta["kernel_ids"].items():
            if name.endswith("_batched"):
                if self.lib.nrpy_set_block_shape(kid, *shape) != 0:
                    raise NrpyError("tile shape %s rejected" % (shape,))
        check(self.lib.nrpy_batch_prepare(self.h), "nrpy_batch_prepare")
        self.free_graph()

    def __del__(self):
        try:
            self.free_graph()
            self.lib.nrpy_batch_free(self.h)
        except Exception:
            pass


def autotune(be, candidates=TILE_CANDIDATES, reps=3):
    """Time MoL_batch_step for each tile shape on the batch (its state is advanced:
    reset it afterwards) and keep the fastest. Returns (best shape, {shape: seconds})."""
    torch = be.torch
    times = {}
    for shape in candidates:
        be.set_tile(shape)
        be.step_eager(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        be.step_eager(reps)
        torch.cuda.synchronize()
        times[shape] = (time.perf_counter() - t0) / reps
    best = min(times, key=times.get)
    be.set_tile(best)
    return best, times
