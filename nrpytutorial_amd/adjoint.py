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
