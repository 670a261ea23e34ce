"""Force matching on the GPU adjoint kernel, wrapped for PyTorch.

Parameters theta = (eps_ab for a <= b, sigma_ab for a <= b, qscale); for the
two-type NaCl model that is (eps++, eps+-, eps--, sig++, sig+-, sig--, qscale).
The engine returns dloss/dlj1[a,b] and dloss/dlj2[a,b] per ordered table
entry (lj1 = 48 eps sig^12, lj2 = 24 eps sig^6, src/simulation.cu:191-195);
the chain rule to (eps, sigma) is applied here.

PyTorch is used only for the fitting driver: FMLoss is a
torch.autograd.Function whose forward calls the C ABI (md_fm_loss_grad) and
whose backward returns the adjoint gradient.  ForceMatch needs an engine
object with fm_set_frames / fm_loss_grad (md_gpu_amd.engine.Engine).
"""
from __future__ import annotations

from typing import Tuple

import numpy as np
import torch

from .systems import System, eps_sig_from_lj


class ForceMatch:
    """Loss and gradient of the force-matching objective for a fixed set of frames."""

    def __init__(self, engine, base: System, frames: np.ndarray, targets: np.ndarray):
        self.e = engine
        self.base = base
        self.nt = base.ff.ntype
        self.iu = np.triu_indices(self.nt)
        engine.fm_set_frames(np.asarray(frames, np.float32), np.asarray(targets, np.float32))

    @property
    def nparam(self) -> int:
        return 2 * len(self.iu[0]) + 1

    def theta0(self) -> np.ndarray:
        eps, sig = eps_sig_from_lj(self.base.ff.lj1, self.base.ff.lj2)
        return np.concatenate([eps[self.iu], sig[self.iu], [1.0]])

    def unpack(self, theta: np.ndarray):
        k = len(self.iu[0])
        E = np.zeros((self.nt, self.nt))
        Sg = np.zeros((self.nt, self.nt))
        E[self.iu] = theta[:k]
        Sg[self.iu] = theta[k:2 * k]
        E[(self.iu[1], self.iu[0])] = theta[:k]
        Sg[(self.iu[1], self.iu[0])] = theta[k:2 * k]
        return E, Sg, float(theta[2 * k])

    def tables(self, theta: np.ndarray):
        E, Sg, q = self.unpack(theta)
        return 48.0 * E * Sg ** 12, 24.0 * E * Sg ** 6, q

    def loss_grad(self, theta: np.ndarray) -> Tuple[float, np.ndarray]:
        theta = np.asarray(theta, np.float64)
        E, Sg, q = self.unpack(theta)
        lj1, lj2, _ = self.tables(theta)
        loss, g1, g2, gq = self.e.fm_loss_grad(lj1, lj2, q)
        # chain rule per ordered entry, then fold (a,b) and (b,a) into the symmetric parameter
        dE = g1 * 48.0 * Sg ** 12 + g2 * 24.0 * Sg ** 6
        dS = g1 * 48.0 * 12.0 * E * Sg ** 11 + g2 * 24.0 * 6.0 * E * Sg ** 5
        dE_sym = dE + dE.T - np.diag(np.diag(dE))
        dS_sym = dS + dS.T - np.diag(np.diag(dS))
        grad = np.concatenate([dE_sym[self.iu], dS_sym[self.iu], [gq]])
        return loss, grad


class FMLoss(torch.autograd.Function):
    """loss(theta) with the GPU adjoint gradient.

    theta: float64 tensor of ForceMatch.nparam entries.  forward calls the C
    ABI once (md_fm_loss_grad: loss and gradient in the same pass); backward
    returns grad_output times that gradient, so no second evaluation is needed.
    """

    @staticmethod
    def forward(ctx, theta, fm):
        loss, grad = fm.loss_grad(theta.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device))
        return torch.as_tensor(loss, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None


def fm_loss(theta, fm: ForceMatch):
    """Differentiable loss for torch: fm_loss(theta_tensor, fm).backward()."""
    return FMLoss.apply(theta, fm)


def fit(fm: ForceMatch, theta0=None, steps: int = 100, lr: float = 1e-3, log_every: int = 0):
    """Adam on theta (relative parameterisation theta = theta0 * (1 + u)); returns (theta, losses)."""
    t0 = torch.as_tensor(fm.theta0() if theta0 is None else theta0, dtype=torch.float64)
    u = torch.zeros_like(t0, requires_grad=True)
    opt = torch.optim.Adam([u], lr=lr)
    losses = []
    for it in range(steps):
        opt.zero_grad()
        loss = fm_loss(t0 * (1.0 + u), fm)
        loss.backward()
        opt.step()
        losses.append(loss.detach().item())
        if log_every and it % log_every == 0:
            print(f"fit {it}: loss {losses[-1]:.6e}", flush=True)
    return (t0 * (1.0 + u)).detach().numpy(), losses


def nacl_problem(engine_cls, nframes: int = 64, nsites: int = 16, seed: int = 9, equil: int = 200, stride: int = 10):
    """The NaCl force-matching problem: frames sampled every `stride` steps of a
    GPU trajectory of the base model (nsites^3 ions, after `equil` steps), and
    target forces of the perturbed model (eps, sigma and the charge scale each
    times 1 + U(-5 %, +5 %), seeded).  Returns (engine, ForceMatch, base,
    frames, targets, theta_target); the engine holds the frames."""
    import dataclasses
    from . import systems as S
    base = S.nacl(nsites=nsites)
    e = engine_cls.from_system(base)
    e.run(equil, stats=False)
    frames = []
    for _ in range(nframes):
        e.run(stride, stats=False)
        frames.append(e.download()[0].copy())
    frames = np.stack(frames)
    tgt_ff, tgt_q = S.perturb_nacl_ff(base.ff, seed=seed)
    t_sys = dataclasses.replace(base.copy(), ff=tgt_ff)
    t_sys.charge = base.charge * tgt_q
    e2 = engine_cls.from_system(t_sys)
    tgt = []
    for f in frames:
        e2.upload(f)
        tgt.append(e2.compute_forces()[0])
    e2.close()
    tgt = np.stack(tgt)
    fm = ForceMatch(e, base, frames, tgt)
    eps, sig = eps_sig_from_lj(tgt_ff.lj1, tgt_ff.lj2)
    theta_t = np.concatenate([eps[fm.iu], sig[fm.iu], [tgt_q]])
    return e, fm, base, frames, tgt, theta_t
