"""Pin the oracle's integrators and force-matching gradient with closed forms,
invariants and finite differences."""
import dataclasses
import math

import numpy as np
import pytest

from md_gpu_amd import systems as S


def gas(n=50, L=200.0, T=300.0, seed=0, integrator="leapfrog", thermostat="none", dt=2.0, tau=100.0):
    """Non-interacting atoms (all LJ tables zero, no charge)."""
    rng = np.random.default_rng(seed)
    ff = S.ForceField(["X"], np.array([20.0]), np.zeros((1, 1)), np.zeros((1, 1)), coulomb=S.COUL_NONE,
                      rc=5.0, skin=1.0)
    pos = rng.uniform(-L / 2, L / 2, (n, 3)).astype(np.float32)
    typ = np.zeros(n, np.int32)
    vel = S.maxwell_boltzmann(rng, ff.masses[typ], T).astype(np.float32)
    return S.System("gas", np.full(3, L, np.float32), pos, vel, typ, np.zeros(n, np.float32), ff,
                    integrator=integrator, thermostat=thermostat, dt=dt, T0=T, tau=tau)


@pytest.mark.parametrize("integ", ["leapfrog", "vverlet"])
def test_free_flight(orc, integ):
    s = gas(integrator=integ)
    r = orc.md_run(s, 25)
    x = s.pos.astype(np.float64) + 25 * s.dt * s.vel.astype(np.float64)
    x -= s.box * np.trunc(2 * x / s.box)   # box centred at 0 (src/nve_leapfrog.cu:50-55)
    d = r["pos"] - x
    d -= s.box * np.rint(d / s.box)
    assert np.abs(d).max() < 1e-9
    assert np.allclose(r["vel"], s.vel, atol=0)


def dimer(k=500.0, r0=1.5, stretch=0.1, integrator="vverlet"):
    ff = S.ForceField(["A", "B"], np.array([12.0, 16.0]), np.zeros((2, 2)), np.zeros((2, 2)), coulomb=S.COUL_NONE,
                      rc=5.0, skin=0.5, bond_k=np.array([k]), bond_r0=np.array([r0]))
    r = r0 + stretch
    pos = np.array([[-r / 2, 0, 0], [r / 2, 0, 0]], np.float32)
    return S.System("dimer", np.full(3, 40.0, np.float32), pos, np.zeros_like(pos), np.array([0, 1], np.int32),
                    np.zeros(2, np.float32), ff, bonds=np.array([[0, 1, 0]], np.int32), integrator=integrator,
                    thermostat="none", dt=0.2)


def test_harmonic_dimer_period(orc):
    """E = k (r - r0)^2 -> spring constant 2k, omega = sqrt(2k/mu) (mu in 1e4 g/mol)."""
    k = 500.0
    s = dimer(k)
    n = 4000
    r = orc.md_run(s, n)
    # bond length series from positions is not returned per step: use the energy oscillation
    pe = r["pe"]
    mu = 1e4 * 12.0 * 16.0 / 28.0
    omega = math.sqrt(2 * k / mu)
    period_energy = math.pi / omega          # PE oscillates at twice the bond frequency
    t = np.arange(n) * s.dt
    pe_c = pe - pe.mean()
    crossings = np.where((pe_c[:-1] < 0) & (pe_c[1:] >= 0))[0]
    measured = np.mean(np.diff(t[crossings]))
    assert measured == pytest.approx(period_energy, rel=5e-3)
    # NVE velocity Verlet conserves the energy to O(dt^2)
    H = r["conserved"]
    assert np.ptp(H) < 1e-3 * (pe.max() - pe.min())


def test_velocity_verlet_time_reversible(orc):
    s = S.argon(ncell=3)
    rng = np.random.default_rng(2)
    s.pos = S.wrap_centered(s.pos + rng.normal(0, 0.1, s.pos.shape), s.box).astype(np.float32)
    s.ff = dataclasses.replace(s.ff, skin=2.0)
    r1 = orc.md_run(s, 40)
    back = s.copy()
    r2 = orc.md_run(back, 40, pos=r1["pos"], vel=-r1["vel"])
    d = r2["pos"] - s.pos.astype(np.float64)
    d -= s.box * np.rint(d / s.box)
    assert np.abs(d).max() < 1e-6


def test_berendsen_relaxation_rate(orc):
    """Free particles with leapfrog Berendsen: T_{n+1} = T_n + (dt/tau)(T0 - T_n) exactly
    (lambda from the previous step's T, src/integrator.cu:146-155)."""
    s = gas(T=150.0, thermostat="berendsen", tau=50.0)
    s.T0 = 300.0
    r = orc.md_run(s, 60)
    T = r["temp"]
    # first step uses T_prev = T0 (lambda = 1), src/integrator.cu:77
    assert T[0] == pytest.approx(s.temperature(), rel=1e-12)
    pred = T[0]
    for n in range(1, 60):
        pred = pred + s.dt / s.tau * (s.T0 - pred)
        assert T[n] == pytest.approx(pred, rel=1e-10)


def test_nose_hoover_conserved_and_mean_temperature(orc):
    s = S.argon(ncell=4)
    rng = np.random.default_rng(3)
    s.pos = S.wrap_centered(s.pos + rng.normal(0, 0.05, s.pos.shape), s.box).astype(np.float32)
    s.thermostat = "nosehoover"
    s.T0 = 120.0
    s.tau = 200.0
    r = orc.md_run(s, 3000)
    H = r["conserved"]
    ke = r["ke"]
    assert np.ptp(H[200:]) < 0.02 * np.ptp(ke[200:] + r["pe"][200:]) + 1e-6 * abs(H[0])
    assert np.mean(r["temp"][1000:]) == pytest.approx(120.0, rel=0.08)


def test_nve_energy_conservation_argon(orc):
    s = S.argon()
    r = orc.md_run(s, 300)
    H = r["conserved"]
    assert np.ptp(H) < 2e-4 * abs(H[0])


# ------------------------------------------------------------------ force matching

def fm_setup(orc, nsites=8, nf=2, seed=5):
    base = S.nacl(nsites=nsites, ff=S.nacl_ff(rc=9.0))
    rng = np.random.default_rng(seed)
    frames = np.stack([S.wrap_centered(base.pos + rng.normal(0, 0.2, base.pos.shape), base.box)
                       for _ in range(nf)]).astype(np.float32)
    tgt_ff, tgt_q = S.perturb_nacl_ff(base.ff, seed=seed)
    targets = []
    q64 = base.charge.astype(np.float64) * tgt_q          # FP64 charges, as the loss uses them
    for f in frames:
        t = base.copy()
        t.pos = f
        rc2, rl2, _, _ = orc.thresholds(tgt_ff.rc, 0.5, t.box)
        counts, nb = orc.nlist(f, t.box, rl2)
        targets.append(orc.pair_forces(t, counts, nb, rc2, charge=q64, ff=tgt_ff, want_energy=False)[0])
    return base, frames, np.stack(targets), tgt_ff, tgt_q


def test_fm_gradient_matches_central_differences(orc):
    base, frames, targets, _, _ = fm_setup(orc)
    ff = base.ff
    loss, g1, g2, gq = orc.fm_loss_grad(frames, base.type, base.charge, base.box, ff, 1.0, targets)

    def L(lj1, lj2, q):
        return orc.fm_loss_grad(frames, base.type, base.charge, base.box, ff, q, targets, lj1, lj2)[0]
    for (a, b) in [(0, 0), (0, 1), (1, 1)]:
        for which, g in ((1, g1), (2, g2)):
            base_t = ff.lj1 if which == 1 else ff.lj2
            h = 1e-4 * base_t[a, b]
            tp = base_t.copy(); tp[a, b] += h
            tm = base_t.copy(); tm[a, b] -= h
            if which == 1:
                fd = (L(tp, ff.lj2, 1.0) - L(tm, ff.lj2, 1.0)) / (2 * h)
            else:
                fd = (L(ff.lj1, tp, 1.0) - L(ff.lj1, tm, 1.0)) / (2 * h)
            assert g[a, b] == pytest.approx(fd, rel=2e-4, abs=1e-12 * abs(loss) / h)
    h = 1e-4
    fdq = (L(ff.lj1, ff.lj2, 1 + h) - L(ff.lj1, ff.lj2, 1 - h)) / (2 * h)
    assert gq == pytest.approx(fdq, rel=2e-5)


def test_fm_zero_gradient_at_target(orc):
    base, frames, targets, tgt_ff, tgt_q = fm_setup(orc)
    loss, g1, g2, gq = orc.fm_loss_grad(frames, base.type, base.charge, base.box, tgt_ff, tgt_q, targets)
    assert loss < 1e-20
    assert np.abs(g1).max() < 1e-12 and np.abs(g2).max() < 1e-12 and abs(gq) < 1e-9
