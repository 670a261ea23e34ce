"""The force-matching driver (md_gpu_amd/fit.py) on the CPU.

ForceMatch turns the engine's gradient with respect to the ordered LJ table
entries (lj1 = 48 eps sig^12, lj2 = 24 eps sig^6, src/simulation.cu:191-195)
and the charge scale into the gradient with respect to theta = (eps_ab,
sig_ab for a <= b, qscale).  Here the engine is a stub whose fm_loss_grad is
the FP64 oracle, so the chain rule, the folding of (a, b) and (b, a) into one
parameter and the torch wrapper are checked against central finite
differences of the loss itself, with no GPU.
"""
import dataclasses

import numpy as np
import pytest
import torch

import oracle
from md_gpu_amd import systems as S
from md_gpu_amd.fit import FMLoss, ForceMatch, fit, fm_loss


class OracleEngine:
    """fm_set_frames / fm_loss_grad with the FP64 oracle behind them."""

    def __init__(self, base):
        self.base = base
        self.calls = 0

    def fm_set_frames(self, frames, targets):
        self.frames = np.asarray(frames, np.float32)
        self.targets = np.asarray(targets, np.float64)

    def fm_loss_grad(self, lj1, lj2, qscale):
        self.calls += 1
        b = self.base
        return oracle.fm_loss_grad(self.frames, b.type, b.charge, b.box, b.ff, qscale, self.targets, lj1, lj2)


def _problem(nframes=2, seed=5):
    base = S.nacl(nsites=10)                      # 1,000 ions, L = 28.2 A > 2 (rc + skin)
    rng = np.random.default_rng(seed)
    frames = np.stack([S.wrap_centered(base.pos + rng.normal(0.0, 0.15, base.pos.shape), base.box[0])
                       for _ in range(nframes)]).astype(np.float32)
    tgt_ff, tgt_q = S.perturb_nacl_ff(base.ff, seed=9)
    t_sys = dataclasses.replace(base.copy(), ff=tgt_ff)
    t_sys.charge = base.charge * tgt_q
    targets = []
    for f in frames:
        t = t_sys.copy()
        t.pos = f
        targets.append(oracle.forces(t)[0])
    eng = OracleEngine(base)
    fm = ForceMatch(eng, base, frames, np.stack(targets))
    eps, sig = S.eps_sig_from_lj(tgt_ff.lj1, tgt_ff.lj2)
    theta_t = np.concatenate([eps[fm.iu], sig[fm.iu], [tgt_q]])
    return fm, eng, theta_t


@pytest.fixture(scope="module")
def problem():
    return _problem()


def test_theta_layout(problem):
    fm, _, _ = problem
    assert fm.nparam == 7
    th = fm.theta0()
    lj1, lj2, q = fm.tables(th)
    np.testing.assert_allclose(lj1, fm.base.ff.lj1, rtol=1e-12)
    np.testing.assert_allclose(lj2, fm.base.ff.lj2, rtol=1e-12)
    assert q == 1.0


def _central(fm, th, k, h):
    tp, tm = th.copy(), th.copy()
    tp[k] += h
    tm[k] -= h
    return (fm.loss_grad(tp)[0] - fm.loss_grad(tm)[0]) / (2 * h)


def test_gradient_matches_central_differences(problem):
    """d loss / d theta from the chain rule vs central FD with h = 1e-4 theta_k.
    The plain central difference is off by O(h^2) terms of relative size
    ~ 220 (h / sigma)^2 = 2e-6 for sigma^12, so the step h and h/2 are combined
    (Richardson), which leaves O(h^4)."""
    fm, _, _ = problem
    th = fm.theta0()
    loss, g = fm.loss_grad(th)
    assert loss > 0
    for k in range(fm.nparam):
        h = 1e-4 * abs(th[k])
        d1, d2 = _central(fm, th, k, h), _central(fm, th, k, h / 2)
        assert abs(g[k] - d1) <= 1e-5 * np.abs(g).max(), (k, g[k], d1)
        fd = (4.0 * d2 - d1) / 3.0
        assert abs(g[k] - fd) <= 1e-8 * np.abs(g).max() + 1e-8 * abs(fd), (k, g[k], fd)


def test_gradient_vanishes_at_target(problem):
    """The targets are the perturbed model's forces: loss and gradient vanish
    there, up to the float32 rounding of the scaled charges the targets were
    made with (relative 6e-8 on the Coulomb forces)."""
    fm, _, theta_t = problem
    loss0, g0 = fm.loss_grad(fm.theta0())
    loss, g = fm.loss_grad(theta_t)
    assert loss < 1e-12 * loss0
    assert np.abs(g).max() < 1e-5 * np.abs(g0).max()


def test_fmloss_backward_is_the_adjoint_gradient(problem):
    fm, eng, _ = problem
    th = torch.tensor(fm.theta0() * (1.0 + 0.01 * np.arange(7)), dtype=torch.float64, requires_grad=True)
    n0 = eng.calls
    loss = FMLoss.apply(th, fm)
    loss.backward()
    assert eng.calls == n0 + 1          # backward does not evaluate again
    ref_loss, ref_g = fm.loss_grad(th.detach().numpy())   # the oracle's threaded sums: same to ~1e-15
    assert loss.detach().item() == pytest.approx(ref_loss, rel=1e-12)
    np.testing.assert_allclose(th.grad.numpy(), ref_g, rtol=1e-12)
    # chained through further torch ops: d(3 L^2)/d theta = 6 L dL/dtheta
    th.grad = None
    (3.0 * fm_loss(th, fm) ** 2).backward()
    np.testing.assert_allclose(th.grad.numpy(), 6.0 * ref_loss * ref_g, rtol=1e-12)


def test_adam_steps_reduce_loss():
    fm, _, _ = _problem(nframes=1, seed=6)
    theta, losses = fit(fm, steps=8, lr=2e-3)
    assert theta.shape == (7,)
    assert losses[-1] < losses[0]
