"""Pin the oracle's pair, bonded and energy arithmetic against closed forms,
finite differences and invariants (the reference ships no test vectors)."""
import dataclasses
import math

import numpy as np
import pytest

from md_gpu_amd import systems as S
from conftest import rel_rms


def two_body(ff, r, q=(0.0, 0.0), types=(0, 0), L=60.0):
    box = np.full(3, L, np.float32)
    pos = np.array([[-r / 2, 0.1, -0.2], [r / 2, 0.1, -0.2]], np.float32)
    s = S.System("pair", box, pos, np.zeros_like(pos), np.array(types, np.int32), np.array(q, np.float32), ff)
    return s


def pair_force_energy(orc, ff, r, q=(0.0, 0.0), erfc_exact=0):
    s = two_body(ff, r, q)
    d = float(s.pos[0, 0]) - float(s.pos[1, 0])       # exact float separation actually used
    counts = np.array([1, 1], np.int32)
    nbrs = np.array([[1], [0]], np.int32)
    rc2 = np.float32(ff.rc) * np.float32(ff.rc)
    F, e, w = orc.pair_forces(s, counts, nbrs, float(rc2), erfc_exact)
    return F, e, w, d


def lj_ff(eps=0.9, sig=3.4, rc=10.0):
    return S.ForceField.from_eps_sig(["A"], [40.0], [[eps]], [[sig]], coulomb=S.COUL_NONE, rc=rc, skin=1.0)


def test_lj_zero_force_at_minimum(orc):
    ff = lj_ff()
    rmin = 2 ** (1 / 6) * 3.4
    F, e, w, d = pair_force_energy(orc, ff, rmin)
    # force scale at r = sigma for reference
    F2, _, _, _ = pair_force_energy(orc, ff, 3.4)
    assert abs(F[0, 0]) < 1e-5 * abs(F2[0, 0])


def test_lj_energy_closed_form_and_shift(orc):
    eps, sig, rc = 0.9, 3.4, 10.0
    ff = lj_ff(eps, sig, rc)
    for r in (3.2, 3.4, 4.1, 7.5):
        _, e, _, d = pair_force_energy(orc, ff, r)
        u = lambda x: 4 * eps * ((sig / x) ** 12 - (sig / x) ** 6)
        assert e[0] == pytest.approx(u(abs(d)) - u(rc), rel=1e-12, abs=1e-14)


# dyadic distances and step: every coordinate is exact in float32
@pytest.mark.parametrize("r", [3.125, 3.8125, 5.0, 8.875])
def test_lj_force_is_minus_gradient(orc, r):
    ff = lj_ff()
    h = 2.0 ** -12
    _, ep, _, dp = pair_force_energy(orc, ff, r + h)
    _, em, _, dm = pair_force_energy(orc, ff, r - h)
    F, _, _, d = pair_force_energy(orc, ff, r)
    fd = -(ep[0] - em[0]) / (abs(dp) - abs(dm))
    # force on atom 0 along +x is -F_radial (atom 0 sits at -r/2)
    assert -F[0, 0] == pytest.approx(fd, rel=1e-5, abs=1e-6)


def wolf_ff(alpha=0.2, rc=12.0, kind=S.COUL_WOLF):
    return S.ForceField.from_eps_sig(["Na", "Cl"], [23.0, 35.45], np.zeros((2, 2)), np.zeros((2, 2)),
                                     coulomb=kind, rc=rc, alpha=alpha, skin=1.0)


@pytest.mark.parametrize("r", [2.5, 4.0, 7.0, 11.0])
def test_wolf_force_is_minus_gradient_exact_erfc(orc, r):
    ff = wolf_ff()
    h = 2.0 ** -12
    q = (1.0, -1.0)
    _, ep, _, dp = pair_force_energy(orc, ff, r + h, q, erfc_exact=1)
    _, em, _, dm = pair_force_energy(orc, ff, r - h, q, erfc_exact=1)
    F, _, _, _ = pair_force_energy(orc, ff, r, q, erfc_exact=1)
    fd = -(ep[1] - em[1]) / (abs(dp) - abs(dm))
    assert -F[0, 0] == pytest.approx(fd, rel=1e-6, abs=1e-6)


def test_wolf_force_vanishes_at_cutoff(orc):
    ff = wolf_ff()
    F1, _, _, _ = pair_force_energy(orc, ff, 11.999, (1.0, -1.0))
    F0, _, _, _ = pair_force_energy(orc, ff, 4.0, (1.0, -1.0))
    assert abs(F1[0, 0]) < 1e-3 * abs(F0[0, 0])
    _, e, _, _ = pair_force_energy(orc, ff, 11.999, (1.0, -1.0))
    assert abs(e[1]) < 1e-6


def test_wolf_alpha_to_zero_limit(orc):
    """alpha -> 0: F = qq (1/r^2 - 1/(rc r)) (shifted-force Coulomb of src/pair_forces.cu:178-183)."""
    ff = wolf_ff(alpha=1e-9)
    for r in (3.0, 6.0, 10.0):
        F, _, _, d = pair_force_energy(orc, ff, r, (1.0, 1.0))
        rr = abs(d)
        expect = S.QFAC * (1 / rr ** 2 - 1 / (12.0 * rr))
        assert -F[0, 0] == pytest.approx(expect, rel=1e-6)


def test_wolf_as_vs_exact_erfc(orc):
    """The A&S 7.1.26 polynomial is within 1.5e-7 of erfc (absolute), so the two
    oracle modes differ by at most ~qfac * 2e-7 / r^2 in force."""
    ff = wolf_ff()
    for r in (2.4, 3.3, 6.0, 9.0):
        Fa, _, _, d = pair_force_energy(orc, ff, r, (1.0, -1.0), 0)
        Fe, _, _, _ = pair_force_energy(orc, ff, r, (1.0, -1.0), 1)
        bound = S.QFAC * 4e-7 / d ** 2
        assert abs(Fa[0, 0] - Fe[0, 0]) <= bound


def test_reaction_field_closed_form(orc):
    eps_rf, rc = 78.0, 12.0
    ff = wolf_ff(rc=rc, kind=S.COUL_RF)
    ff = dataclasses.replace(ff, eps_rf=eps_rf)
    kr = 2 * (eps_rf - 1) / (2 * eps_rf + 1) / rc ** 3
    for r in (3.0, 7.0, 11.5):
        F, e, _, d = pair_force_energy(orc, ff, r, (1.0, -1.0))
        rr = abs(d)
        fr = -S.QFAC * (1 / rr ** 2 - kr * rr)          # radial force, qq = -1
        assert -F[0, 0] == pytest.approx(fr, rel=1e-12)
        crf = 1 / rc + 0.5 * kr * rc ** 2
        assert e[1] == pytest.approx(-S.QFAC * (1 / rr + 0.5 * kr * rr ** 2 - crf), rel=1e-12)


def perturbed(sys_, amp, seed=1):
    s = sys_.copy()
    rng = np.random.default_rng(seed)
    s.pos = S.wrap_centered(s.pos + rng.normal(0, amp, s.pos.shape), s.box).astype(np.float32)
    return s


@pytest.mark.parametrize("mk", [lambda: perturbed(S.argon(ncell=4), 0.2),
                                lambda: perturbed(S.nacl(nsites=10), 0.2),
                                lambda: perturbed(S.water(nmol=216), 0.05)])
def test_total_force_vanishes(orc, mk):
    s = mk()
    F, e, w = orc.forces(s)
    assert np.abs(F.sum(0)).max() <= 1e-9 * np.abs(F).sum()


def test_virial_is_scaling_derivative(orc):
    """W = sum r.F = -dU/dlambda for x -> lambda x, L -> lambda L (pairs only)."""
    s = perturbed(S.nacl(nsites=8, ff=S.nacl_ff(rc=9.0)), 0.15)
    s.ff = dataclasses.replace(s.ff, skin=0.5)

    def U(lam):
        t = s.copy()
        t.pos = (s.pos.astype(np.float64) * lam).astype(np.float32)
        t.box = (s.box.astype(np.float64) * lam).astype(np.float32)
        # scale in float64, the float32 rounding is the same for both signs of h
        F, e, w = orc.forces(t, erfc_exact=1)
        return e[0] + e[1], w
    h = 2e-4
    up, _ = U(1 + h)
    um, _ = U(1 - h)
    _, w = U(1.0)
    assert w == pytest.approx(-(up - um) / (2 * h), rel=2e-3)


def test_bond_force_is_minus_gradient(orc):
    s = S.water(nmol=1)
    s.box = np.full(3, 40.0, np.float32)
    s.ff = dataclasses.replace(s.ff, coulomb=S.COUL_NONE, lj1=np.zeros((2, 2)), lj2=np.zeros((2, 2)))
    rng = np.random.default_rng(4)
    s.pos = (s.pos + rng.normal(0, 0.05, s.pos.shape)).astype(np.float32)
    F, e, _ = orc.bonded_forces(s)
    h = 1e-3
    for a in range(3):
        for d in range(3):
            p1 = s.pos.astype(np.float64).copy(); p1[a, d] += h
            p2 = s.pos.astype(np.float64).copy(); p2[a, d] -= h
            _, e1, _ = orc.bonded_forces(s, p1.astype(np.float32))
            _, e2, _ = orc.bonded_forces(s, p2.astype(np.float32))
            dx = float(np.float32(p1[a, d])) - float(np.float32(p2[a, d]))
            fd = -((e1[0] + e1[1]) - (e2[0] + e2[1])) / dx
            assert F[a, d] == pytest.approx(fd, rel=2e-3, abs=2e-2)
    # sum of forces and torque vanish for an isolated molecule
    assert np.abs(F.sum(0)).max() < 1e-9
    tq = np.cross(s.pos.astype(np.float64), F).sum(0)
    assert np.abs(tq).max() < 1e-8 * np.abs(F).max() * 10


def test_angle_energy_closed_form(orc):
    s = S.water(nmol=1)
    s.box = np.full(3, 40.0, np.float32)
    p = s.pos.astype(np.float64)
    a, b = p[1] - p[0], p[2] - p[0]
    th = math.acos(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b))
    _, e, _ = orc.bonded_forces(s)
    assert e[1] == pytest.approx(383.0 * (th - math.radians(109.47)) ** 2, abs=1e-9)
    r1 = np.linalg.norm(a)
    r2 = np.linalg.norm(b)
    assert e[0] == pytest.approx(4637.0 * ((r1 - 1) ** 2 + (r2 - 1) ** 2), abs=1e-8)
