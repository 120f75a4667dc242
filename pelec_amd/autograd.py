"""torch.autograd.Function over the gfx950 BDF reactor: the end state (rhoY, T) as a function of the
initial rhoY, rho*e and the per-reaction log pre-exponential factors, with the hand-written
discrete adjoint kernel (prc_adjoint) as the backward. The forward and the backward each run the
integration on the device; the backward recomputes the forward from checkpoints.

    from pelec_amd.autograd import react, initial_state
    rY0, rE0 = initial_state("drm19", T0, Y0, rho)       # differentiable in T0, Y0
    rY, T = react(reactor, rY0, T0, rE0, logA, dt)
    loss = (w * T).sum() + (v * rY[oh] / rY.sum(0)).sum()
    loss.backward()                                      # grads of logA, T0, Y0
"""
from __future__ import annotations

import numpy as np
import torch

from . import mechanism as M

R_ERG = 8.31446261815324e7


class _React(torch.autograd.Function):
    @staticmethod
    def forward(ctx, reactor, rY0, T0, rE0, logA, dt):
        rY0c, T0c, rE0c = rY0.detach().contiguous(), T0.detach().contiguous(), rE0.detach().contiguous()
        la = None if logA is None else logA.detach().contiguous()
        out = reactor.adjoint(rY0c, T0c, rE0c, dt, logA=la)
        if int((out["status"] != 0).sum()) != 0:
            raise RuntimeError("pelec_amd: integration failed in %d cells"
                               % int((out["status"] != 0).sum()))
        ctx.reactor, ctx.dt = reactor, dt
        ctx.has_logA = logA is not None
        ctx.save_for_backward(rY0c, T0c, rE0c, la if la is not None else torch.empty(0))
        return out["rY"], out["T"]

    @staticmethod
    def backward(ctx, g_rY, g_T):
        rY0, T0, rE0, la = ctx.saved_tensors
        ct_rY = None if g_rY is None else g_rY.contiguous()
        ct_T = None if g_T is None else g_T.contiguous()
        if ct_rY is None and ct_T is None:
            return None, None, None, None, None, None
        if ct_rY is None:
            ct_rY = torch.zeros_like(rY0)
        if ct_T is None:
            ct_T = torch.zeros_like(T0)
        out = ctx.reactor.adjoint(rY0, T0, rE0, ctx.dt, logA=la if ctx.has_logA else None,
                                  ct_rY=ct_rY, ct_T=ct_T)
        g_logA = out["g_logA"].sum(1) if ctx.has_logA else None
        # T0 is only the Newton guess of T(e, Y): no gradient flows through it
        return None, out["g_rY0"], torch.zeros_like(T0), out["g_rE0"], g_logA, None


def react(reactor, rY0, T0, rE0, logA, dt: float):
    """(rY_end (N, ncells), T_end (ncells,)) = BDF reactor over dt, differentiable in rY0, rE0 and
    logA (per-reaction log multipliers of the pre-exponential factors; a (NRXN,) tensor or None)."""
    return _React.apply(reactor, rY0, T0, rE0, logA, dt)


def initial_state(mech: str, T0, Y0, rho):
    """rhoY0 = rho Y0 and rho e0 = rho sum_k Y0_k e_k(T0) (NASA-7, CGS) as torch expressions, so
    gradients reach T0 and Y0 (Y0: (N, ncells); T0, rho: (ncells,))."""
    m = M.load(mech)
    dev, dt_ = T0.device, torch.float64
    low = torch.tensor(np.array([s.low for s in m.species]), dtype=dt_, device=dev)    # (N, 7)
    high = torch.tensor(np.array([s.high for s in m.species]), dtype=dt_, device=dev)
    tmid = torch.tensor([s.Tmid for s in m.species], dtype=dt_, device=dev)[:, None]
    W = torch.tensor([s.W for s in m.species], dtype=dt_, device=dev)[:, None]
    T = T0[None, :]
    a = torch.where((T < tmid)[..., None], low[:, None, :], high[:, None, :])         # (N, nc, 7)
    hRT = (a[..., 0] + T * (a[..., 1] / 2 + T * (a[..., 2] / 3 + T * (a[..., 3] / 4 + T * a[..., 4] / 5)))
           + a[..., 5] / T)
    e = (hRT - 1.0) * R_ERG * T / W                                                  # (N, nc)
    return rho[None, :] * Y0, rho * (Y0 * e).sum(0)
