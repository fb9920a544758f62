"""Reverse-mode differentiation of method-of-lines evolutions.

The forward RK4 step and its adjoint are both compiled HIP code in the bundle:
``rhs_eval_vjp`` (SymPy-derived pointwise adjoint + transposed stencils),
``bc_apply_T`` (transpose of the ghost fill) and ``MoL_step_adjoint``, which
recomputes the three stage states of one step from its starting state and
runs the four stage adjoints (see the generated code for the algebra). This
module adds what lies above one step:

* ``forward_checkpointed``: N steps keeping every K-th state (K ~ sqrt(N));
* ``backward_checkpointed``: for each segment, last first, recompute its
  states from the checkpoint, then step the cotangent back through them, so
  memory holds ~2 sqrt(N) states and each forward step is run twice;
* ``RK4Evolve``: a ``torch.autograd.Function`` wrapping both, so a loss
  written in torch on the final state differentiates through the evolution.

The evolution's output depends only on the interior of its input (the ghost
zones are refilled by the boundary conditions first), so the cotangent
returned for the input is zero at ghost points.
"""
import ctypes
import math

from .runtime import NrpyError, check, ptr

_VP = ctypes.c_void_p


class RK4Adjoint:
    """Scratch arrays and entry points for adjoint RK4 steps on one Evolution."""

    def __init__(self, ev):
        b = ev.g.bundle
        if "vjp_slots" not in b.meta:
            raise NrpyError("bundle %s has no generated adjoint" % b.name)
        self.ev = ev
        self.lib = b.lib
        g = ev.g
        n = len(b.evol)
        self.U = [g.empty(n) for _ in range(3)]
        self.rt, self.acc, self.kbar, self.w = (g.empty(n) for _ in range(4))
        self.G = g.empty(len(b.meta["vjp_slots"]))
        f = self.lib.MoL_step_adjoint
        f.restype = ctypes.c_int
        f.argtypes = [_VP] * 14 + [ctypes.c_double, _VP]
        f = self.lib.bc_apply_T
        f.restype = ctypes.c_int
        f.argtypes = [_VP, _VP, ctypes.c_int, _VP, _VP, _VP]

    def step_back(self, y_n, lam, dt):
        """lam <- (d y_{n+1} / d y_n)^T lam for the RK4 step of size dt from y_n."""
        ev = self.ev
        check(self.lib.MoL_step_adjoint(ev.g.rfm, ev.g.pp(), ev.g.bc, ptr(y_n), ptr(lam), ptr(self.U[0]),
                                        ptr(self.U[1]), ptr(self.U[2]), ptr(self.rt), ptr(self.acc), ptr(self.kbar),
                                        ptr(self.w), ptr(self.G), ptr(ev.auxevol), float(dt), ev.s()),
              "MoL_step_adjoint")

    def bc_T(self, cot):
        """cot <- E^T cot: ghost cotangents folded onto their interior sources, then zeroed."""
        ev = self.ev
        check(self.lib.bc_apply_T(ev.g.pp(), ev.g.bc, len(ev.g.bundle.evol), ev.g.bundle.evol_parity, ptr(cot),
                                  ev.s()), "bc_apply_T")


def segment_length(nsteps):
    return max(1, int(math.ceil(math.sqrt(nsteps))))


def forward_checkpointed(ev, y0, dt, nsteps, K=None):
    """Evolve y0 (its interior; ghosts are refilled) for nsteps RK4 steps on ev.
    Returns (final state tensor, {step: state} checkpoints every K steps)."""
    K = K or segment_length(nsteps)
    ev.y_n.copy_(y0)
    ev.apply_bcs(ev.y_n, composed=True)
    ev.post_bcs(ev.y_n)
    cps = {0: ev.y_n.clone()}
    done = 0
    while done < nsteps:
        m = min(K, nsteps - done)
        ev.step(dt, m, use_graph=True)
        done += m
        if done < nsteps:
            cps[done] = ev.y_n.clone()
    return ev.y_n.clone(), cps, K


def backward_checkpointed(ev, adj, cps, K, dt, nsteps, lam_T):
    """Cotangent of the initial state from the cotangent lam_T of the final state."""
    lam = lam_T.clone()
    adj.bc_T(lam)
    for start in sorted(cps, reverse=True):
        end = min(start + K, nsteps)
        ev.y_n.copy_(cps[start])
        states = [cps[start]]
        for _ in range(start, end - 1):
            ev.step(dt, 1, use_graph=True)
            states.append(ev.y_n.clone())
        for s in reversed(states):
            adj.step_back(s, lam, dt)
    return lam


def _autograd_function():
    import torch

    class RK4Evolve(torch.autograd.Function):
        """y_T = RK4^nsteps(y0) on the GPU, differentiable in y0 (sqrt(N) checkpointing).

        ``RK4Evolve.apply(y0, ev, adj, dt, nsteps)``: y0 a state tensor of ev's grid shape."""

        @staticmethod
        def forward(ctx, y0, ev, adj, dt, nsteps):
            yT, cps, K = forward_checkpointed(ev, y0.detach(), dt, nsteps)
            ctx.saved = (ev, adj, cps, K, dt, nsteps)
            return yT

        @staticmethod
        def backward(ctx, gT):
            ev, adj, cps, K, dt, nsteps = ctx.saved
            lam0 = backward_checkpointed(ev, adj, cps, K, dt, nsteps, gT.contiguous())
            return lam0, None, None, None, None

    return RK4Evolve


_RK4Evolve = None


def RK4Evolve():
    global _RK4Evolve
    if _RK4Evolve is None:
        _RK4Evolve = _autograd_function()
    return _RK4Evolve


def torch_mul(x, a, out):
    """out = a * x in one device pass."""
    import torch
    torch.mul(x, a, out=out)


class BSSNStepAdjoint:
    """Adjoint of the BSSN RK4 step on one Evolution, orchestrated from the host over the
    bundle's kernels. Each substep of the forward step is
        X_s = D(E(u_s)),  k_{s+1} = R(X_s, Ric(X_s)),  u_{s+1} = y_n + c_s dt k_{s+1},
    with E the ghost fill, D the det(gammabar) enforcement (pointwise at every point),
    Ric the Ricci tensor and R the right-hand side; the step ends with
    y_{n+1} = D(E(y_n + dt/6 (k1 + 2 k2 + 2 k3 + k4))). The backward pass recomputes the
    stage states X_s and the pre-enforcement states P_s = E(u_s), then applies the
    transposes in reverse: D^T (detg VJP at P_s), E^T (bc_apply_T), and
    R^T + Ric^T (rhs_eval_vjp, whose RbarDD cotangent feeds Ricci_eval_vjp)."""

    def __init__(self, ev):
        b = ev.g.bundle
        for key in ("vjp_slots", "ricci_vjp_slots", "detg_vjp_slots"):
            if key not in b.meta:
                raise NrpyError("bundle %s has no %s" % (b.name, key))
        self.ev = ev
        self.lib = b.lib
        g = ev.g
        n = len(b.evol)
        self.n = n
        self.X = [g.empty(n) for _ in range(4)]          # X_0 = y_n copy, X_1..X_3
        self.P = [g.empty(n) for _ in range(5)]          # P_1..P_4 (index 1..4)
        self.Ric = [g.empty(6) for _ in range(4)]
        self.rt, self.yw = g.empty(n), g.empty(n)
        self.acc, self.c, self.cX, self.tmp = (g.empty(n) for _ in range(4))
        self.ck = [g.empty(n) for _ in range(4)]
        self.cRic = g.empty(6)
        nslots = max(len(b.meta["vjp_slots"]), len(b.meta["ricci_vjp_slots"]), len(b.meta["detg_vjp_slots"]),
                     len(b.meta.get("hamiltonian_vjp_slots", [])))
        self.G = g.empty(nslots)
        L = self.lib
        for f in ("rhs_eval_vjp", "Ricci_eval_vjp", "enforce_detgammabar_constraint_vjp", "Hamiltonian_constraint_vjp"):
            fn = getattr(L, f)
            fn.restype = ctypes.c_int
            fn.argtypes = [_VP] * 9
        # index ranges of the Ricci tensor's inputs (hDD, lambdaU) among the evolved gridfunctions
        idx = sorted(b.evol.index(n) for n in b.evol if n.startswith("hDD") or n.startswith("lambdaU"))
        self.ricci_in = []
        for i in idx:
            if self.ricci_in and self.ricci_in[-1][1] == i:
                self.ricci_in[-1] = (self.ricci_in[-1][0], i + 1)
            else:
                self.ricci_in.append((i, i + 1))
        L.bc_apply_T.restype = ctypes.c_int
        L.bc_apply_T.argtypes = [_VP, _VP, ctypes.c_int, _VP, _VP, _VP]
        L.rhs_eval_rk4_substep.restype = ctypes.c_int

    def _E(self, y):
        self.ev.apply_bcs(y, composed=True)

    def _ET(self, c):
        ev = self.ev
        check(self.lib.bc_apply_T(ev.g.pp(), ev.g.bc, self.n, ev.g.bundle.evol_parity, ptr(c), ev.s()), "bc_apply_T")

    def _DT(self, P, c):
        """c <- D'(P)^T c (the enforcement changes only hDD; other cotangents pass through).
        In place: pass 1 of the VJP reads every cotangent it needs into its slots before
        pass 2 writes the hDD cotangents, and the two run in order on the stream."""
        ev = self.ev
        check(self.lib.enforce_detgammabar_constraint_vjp(ev.g.rfm, ev.g.pp(), None, ptr(P), ptr(c), ptr(self.G),
                                                          ptr(c), None, ev.s()), "detg vjp")

    def _RT(self, X, Ric, ck, out):
        """out <- (R + Ric)'(X)^T ck: full-array cotangent of the stage input X."""
        ev = self.ev
        g = ev.g
        check(self.lib.rhs_eval_vjp(g.rfm, g.pp(), ptr(Ric), ptr(X), ptr(ck), ptr(self.G), ptr(out), ptr(self.cRic),
                                    ev.s()), "rhs_eval_vjp")
        # the Ricci VJP writes every point of the cotangents of hDD and lambdaU only
        check(self.lib.Ricci_eval_vjp(g.rfm, g.pp(), None, ptr(X), ptr(self.cRic), ptr(self.G), ptr(self.tmp), None,
                                      ev.s()), "Ricci_eval_vjp")
        for a, b in self.ricci_in:
            out[a:b].add_(self.tmp[a:b])

    def recompute(self, y_n, dt):
        ev, g, L = self.ev, self.ev.g, self.lib
        self.X[0].copy_(y_n)
        self.yw.copy_(y_n)
        for s in range(4):
            X = self.X[s]
            check(L.Ricci_eval(g.rfm, g.pp(), ptr(X), ptr(self.Ric[s]), ev.s()), "Ricci_eval")
            out = self.yw if s == 3 else self.X[s + 1]
            check(L.rhs_eval_rk4_substep(s, g.rfm, g.pp(), ptr(self.Ric[s]), ptr(X), ptr(self.yw), ptr(self.rt),
                                         ptr(out), float(dt), ev.s()), "rhs_eval_rk4_substep")
            self._E(out)
            self.P[s + 1].copy_(out)
            ev.post_bcs(out)

    def step_back(self, y_n, lam, dt):
        """lam (cotangent of y_{n+1}, full array) <- cotangent of y_n (full array)."""
        self.recompute(y_n, dt)
        c = self.c
        c.copy_(lam)
        self._DT(self.P[4], c)
        self._ET(c)                                  # cotangent of u4 (interior)
        self.acc.copy_(c)
        w = (dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0)
        for s in range(4):
            torch_mul(c, w[s], self.ck[s])
        coef = (dt / 2.0, dt / 2.0, dt)               # u_{s+1} = y_n + coef[s] * dt-weighted k_{s+1}
        for s in (3, 2, 1):                         # stage input X_s -> k_{s+1}
            self._RT(self.X[s], self.Ric[s], self.ck[s], self.cX)
            self._DT(self.P[s], self.cX)
            self._ET(self.cX)                        # cotangent of u_s
            self.acc.add_(self.cX)
            self.ck[s - 1].add_(self.cX, alpha=coef[s - 1])
        self._RT(self.X[0], self.Ric[0], self.ck[0], self.cX)
        lam.copy_(self.cX).add_(self.acc)


def bssn_hl2_mass_gradient(ev, nsteps, CFL=0.5):
    """J = sqrt(mean_interior H^2) after nsteps RK4 steps of Brill-Lindquist data on ev,
    and dJ/d(BH1_mass, BH2_mass) by the adjoint (all states of the run are kept: the
    Brill-Lindquist grids this is meant for are small)."""
    import torch
    g, L = ev.g, ev.lib
    adj = BSSNStepAdjoint(ev)
    dt = g.find_timestep(CFL)
    n = len(g.bundle.evol)
    # initial data, its mass derivatives, y_0 = D(E(ID))
    ID = g.empty(n)
    check(L.initial_data(g.pp(), g.xx_dev_ptrs, ptr(ID), ev.s()), "initial_data")
    dm1, dm2 = g.empty(n), g.empty(n)
    L.initial_data_dmass.restype = ctypes.c_int
    check(L.initial_data_dmass(g.pp(), g.xx_dev_ptrs, ptr(dm1), ptr(dm2), ev.s()), "initial_data_dmass")
    P0 = ID.clone()
    ev.apply_bcs(P0, composed=True)
    ev.y_n.copy_(P0)
    ev.post_bcs(ev.y_n)
    states = [ev.y_n.clone()]
    for _ in range(nsteps):
        ev.step(dt, 1, use_graph=True)
        states.append(ev.y_n.clone())
    # loss and its cotangent through the Hamiltonian constraint
    H = ev.hamiltonian(states[-1])
    Hi = g.interior(H)
    J = torch.sqrt(torch.mean(Hi * Hi))
    cH = torch.zeros_like(ev.aux)
    g.interior(cH[0]).copy_(Hi / (Hi.numel() * J))
    lam = torch.zeros_like(ev.y_n)
    check(L.Hamiltonian_constraint_vjp(g.rfm, g.pp(), None, ptr(states[-1]), ptr(cH), ptr(adj.G), ptr(lam), None,
                                       ev.s()), "Hamiltonian_constraint_vjp")
    for s in range(nsteps - 1, -1, -1):
        adj.step_back(states[s], lam, dt)
    adj._DT(P0, lam)
    adj._ET(lam)                                     # cotangent of the interior initial data
    grad = [float((lam * dm).sum()) for dm in (dm1, dm2)]
    ev.free_graph()
    return float(J), grad
