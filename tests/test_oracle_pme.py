"""The oracle's Ewald electrostatics (oracle/md_pme.cpp): smooth particle-mesh
Ewald (Essmann et al., J. Chem. Phys. 103, 8577 (1995)), restated from the paper
because the reference vendors GROMACS' PME (src/pme.c) without building it and
its lkspace 2 path has only the damped real-space term (src/forces.cu:97-110).
Pinned by what the mathematics fixes:

- the oracle's FFT equals numpy.fft on odd, even and prime lengths;
- the B-splines are a partition of unity with derivatives summing to zero, and
  match the closed form of the cubic B-spline;
- SPME converges to the direct Ewald sum over reciprocal vectors as the grid and
  order grow (energy, forces, virial);
- real + reciprocal + self energy of the rock-salt lattice gives the Madelung
  constant 1.747564594633;
- the total Ewald energy does not depend on the splitting parameter beta, with and
  without excluded pairs (a wrong prefactor in any of the four terms breaks this);
- forces are minus the gradient of the energy and the virial is -L dE/dL, by
  central finite differences."""
import dataclasses
import math

import numpy as np
import pytest

import oracle as orc
from md_gpu_amd import systems as S

MADELUNG_NACL = 1.747564594633182


def test_fft_matches_numpy():
    rng = np.random.default_rng(0)
    for K in [(8, 6, 5), (12, 9, 10), (7, 11, 13), (16, 16, 16), (1, 4, 3)]:
        a = rng.normal(size=K) + 1j * rng.normal(size=K)
        f = np.fft.fftn(a)
        assert np.abs(orc.fft3(a, -1) - f).max() < 1e-13 * np.abs(f).max()
        b = np.fft.ifftn(a) * np.prod(K)
        assert np.abs(orc.fft3(a, +1) - b).max() < 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("order", [2, 3, 4, 5, 6, 8, 10])
def test_bspline_partition_of_unity(order):
    for w in (0.0, 0.25, 0.5, 0.9, 0.999):
        th, dth = orc.bspline(order, w)
        assert abs(th.sum() - 1.0) < 1e-14
        assert abs(dth.sum()) < 1e-14
        assert np.all(th >= 0.0)
        # derivative by central differences of the spline itself
        h = 1e-6
        if 0.0 < w - h and w + h < 1.0:
            tp, _ = orc.bspline(order, w + h)
            tm, _ = orc.bspline(order, w - h)
            assert np.abs((tp - tm) / (2 * h) - dth).max() < 1e-8


def test_cubic_bspline_closed_form():
    """M_4(x) = x^3/6 on [0,1], (-3x^3 + 12x^2 - 12x + 4)/6 on [1,2], symmetric."""
    def m4(x):
        if 0 <= x <= 1:
            return x ** 3 / 6
        if 1 <= x <= 2:
            return (-3 * x ** 3 + 12 * x ** 2 - 12 * x + 4) / 6
        if 2 <= x <= 4:
            return m4(4 - x)
        return 0.0
    for w in (0.0, 0.3, 0.77):
        th, _ = orc.bspline(4, w)
        assert np.allclose(th, [m4(w + j) for j in range(4)], atol=1e-15)


def random_neutral(n, L, seed):
    rng = np.random.default_rng(seed)
    x = (rng.random((n, 3)) - 0.5) * np.asarray(L, np.float64)
    q = rng.choice([-1.0, 1.0], n) * rng.uniform(0.5, 1.5, n)
    q -= q.mean()
    return x, q


def test_spme_converges_to_ewald_sum():
    L = np.array([15.0, 16.0, 17.0], np.float32)
    x, q = random_neutral(40, L, 1)
    beta = 0.35
    Ee, Fe, We = orc.ewald_recip(x, q, L, beta, 1.0, kmax=14)
    errs = []
    for K, p in [((16, 16, 18), 4), ((24, 25, 27), 6), ((32, 32, 36), 8), ((48, 48, 50), 10)]:
        E, F, W = orc.pme_recip(x, q, L, K, p, beta, 1.0)
        errs.append((abs(E - Ee) / abs(Ee), np.abs(F - Fe).max() / np.abs(Fe).max(), abs(W - We) / abs(We)))
    for a, b in zip(errs, errs[1:]):
        assert all(y < 0.05 * x for x, y in zip(a, b))   # each refinement gains well over an order
    assert all(e < 1e-10 for e in errs[-1])
    assert errs[0][1] < 5e-3   # order 4 at about 1 A spacing


def rocksalt(nsites, a=5.64):
    s = S.nacl(nsites=nsites)
    return s.pos.astype(np.float64), s.charge.astype(np.float64), s.box, a / 2.0


def real_space(x, q, box, beta, rc, excl=None):
    """Damped real-space energy and forces with the exact erfc (brute force over
    minimum images; rc below half the box)."""
    n = x.shape[0]
    L = np.asarray(box, np.float64)
    E = 0.0
    F = np.zeros_like(x)
    for i in range(n):
        d = x[i] - x
        d -= L * np.rint(d / L)
        r = np.sqrt((d ** 2).sum(1))
        m = (r < rc) & (np.arange(n) != i)
        if excl is not None:
            for j in excl[i]:
                if j >= 0:
                    m[j] = False
        rr = r[m]
        qq = q[i] * q[m]
        E += 0.5 * np.sum(qq * np.array([math.erfc(beta * t) for t in rr]) / rr)
        fr = qq * (np.array([math.erfc(beta * t) for t in rr]) / rr + 2 * beta / math.sqrt(math.pi) * np.exp(-(beta * rr) ** 2)) / rr ** 2
        F[i] = (fr[:, None] * d[m]).sum(0)
    return E, F


def test_madelung_constant():
    x, q, box, r0 = rocksalt(8)       # 512 ions, L = 22.56 A
    beta, rc = 0.36, 11.0             # erfc(beta rc) = 2e-8
    Er, Fr = real_space(x, q, box, beta, rc)
    Ek, Fk, _ = orc.pme_recip(x, q, box, (32, 32, 32), 10, beta, 1.0)
    Es, _ = orc.ewald_self(q, box, beta, 1.0)
    E = Er + Ek + Es
    M = -E / (x.shape[0] / 2) * r0
    assert abs(M - MADELUNG_NACL) < 2e-7, M
    # every ion of the perfect lattice is at a centre of symmetry: no force
    assert np.abs(Fr + Fk).max() < 1e-6


def test_total_energy_independent_of_beta():
    L = np.array([18.0, 18.0, 18.0], np.float32)
    x, q = random_neutral(120, L, 2)
    tot = []
    for beta in (0.42, 0.5, 0.6):
        Er, Fr = real_space(x, q, L, beta, 8.99)
        Ek, Fk, _ = orc.pme_recip(x, q, L, (40, 40, 40), 10, beta, 1.0)
        Es, _ = orc.ewald_self(q, L, beta, 1.0)
        tot.append((Er + Ek + Es, Fr + Fk))
    for E, F in tot[1:]:
        assert abs(E - tot[0][0]) < 2e-7 * abs(tot[0][0])
        assert np.abs(F - tot[0][1]).max() < 2e-6 * np.abs(tot[0][1]).max()


def test_total_energy_independent_of_beta_with_exclusions():
    """Rigid water: the real-space sum leaves out the intramolecular pairs, the
    correction -qq erf(beta r)/r removes them from the reciprocal sum."""
    s = S.water(nmol=216, seed=5, rigid=True)
    x = s.pos.astype(np.float64)
    q = s.charge.astype(np.float64)
    ex = s.exclusions()
    tot = []
    for beta in (0.42, 0.55):
        Er, Fr = real_space(x, q, s.box, beta, 9.0, ex)
        Ek, Fk, _ = orc.pme_recip(x, q, s.box, (40, 40, 40), 10, beta, 1.0)
        Ex, Fx, _ = orc.pme_excl(x, q, s.box, ex, beta, 1.0)
        Es, _ = orc.ewald_self(q, s.box, beta, 1.0)
        tot.append((Er + Ek + Ex + Es, Fr + Fk + Fx))
        scale = abs(Es)   # the terms are tens of times the (neutral molecules') total
    assert abs(tot[1][0] - tot[0][0]) < 2e-8 * scale
    assert np.abs(tot[1][1] - tot[0][1]).max() < 2e-6 * np.abs(tot[0][1]).max()


@pytest.mark.parametrize("order", [4, 6])
def test_forces_are_minus_gradient(order):
    L = np.array([14.0, 15.0, 16.0], np.float32)
    x, q = random_neutral(30, L, 3)
    K, beta = (14, 15, 18), 0.3
    _, F, _ = orc.pme_recip(x, q, L, K, order, beta, 1.0)
    h = 1e-5
    for i in (0, 7, 29):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            g = (orc.pme_recip(xp, q, L, K, order, beta, 1.0)[0] - orc.pme_recip(xm, q, L, K, order, beta, 1.0)[0]) / (2 * h)
            assert abs(-g - F[i, d]) < 1e-7 * np.abs(F).max()


def test_exclusion_forces_are_minus_gradient():
    s = S.water(nmol=8, seed=1, rigid=True)
    x = s.pos.astype(np.float64)
    q = s.charge.astype(np.float64)
    ex = s.exclusions()
    _, F, _ = orc.pme_excl(x, q, s.box, ex, 0.35, 1.0)
    h = 1e-6
    for i in (0, 1, 2, 13):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            g = (orc.pme_excl(xp, q, s.box, ex, 0.35, 1.0)[0] - orc.pme_excl(xm, q, s.box, ex, 0.35, 1.0)[0]) / (2 * h)
            assert abs(-g - F[i, d]) < 1e-7 * np.abs(F).max()


def test_virial_is_minus_L_dE_dL():
    """Uniform scaling of box and positions (box edges exact in float32)."""
    L = np.array([16.0, 16.0, 16.0], np.float32)
    x, q = random_neutral(30, L, 4)
    K, p, beta = (16, 16, 16), 6, 0.33
    _, _, W = orc.pme_recip(x, q, L, K, p, beta, 1.0)
    _, _, We = orc.ewald_recip(x, q, L, beta, 1.0, 12)
    eps = 2.0 ** -10
    E = []
    for s in (1 + eps, 1 - eps, 1 + 2 * eps, 1 - 2 * eps):
        Ls = (L * np.float32(s)).astype(np.float32)
        assert Ls[0] == np.float64(L[0]) * s
        E.append(orc.pme_recip(x * s, q, Ls, K, p, beta, 1.0)[0])
    # fourth-order central difference of E(s) at s = 1
    dE = (8 * (E[0] - E[1]) - (E[2] - E[3])) / (12 * eps)
    assert abs(-dE - W) < 1e-7 * abs(W)
    assert abs(W - We) < 1e-3 * abs(We)
    # self and background: -L d/dL of -pi Q^2 / (2 V beta^2) is 3 times itself
    qn = q + 0.1
    e0, w0 = orc.ewald_self(qn, L, beta, 1.0)
    e1, _ = orc.ewald_self(qn, (L * np.float32(1 + eps)).astype(np.float32), beta, 1.0)
    e2, _ = orc.ewald_self(qn, (L * np.float32(1 - eps)).astype(np.float32), beta, 1.0)
    assert abs(-(e1 - e2) / (2 * eps) - w0) < 1e-5 * abs(w0)


def test_md_with_pme_conserves_energy():
    """NaCl 1,728 ions with damped real space + SPME (order 4, 1 A grid), NVE
    velocity Verlet 1 fs from 600 K: over 400 steps the total energy drifts by
    < 0.2 % of the kinetic energy and stays within 0.5 % of it."""
    s = S.with_pme(S.nacl(nsites=12, T=600.0), spacing=1.0, order=4)
    s = dataclasses.replace(s, thermostat="none", integrator="vverlet", dt=1.0)
    n = 400
    r = orc.md_run(s, n)
    H = r["conserved"]
    ke = r["ke"].mean()
    assert np.all(np.isfinite(H))
    slope = np.polyfit(np.arange(n), H, 1)[0]
    assert abs(slope) * n < 0.002 * ke, (slope * n, ke)
    assert H.max() - H.min() < 0.005 * ke, (H.max() - H.min(), ke)


def test_md_forces_include_long_range():
    """The trajectory's first potential energy with PME minus the one without is
    reciprocal + exclusion + self from the separate functions; runs without the
    damped term are refused."""
    s = S.with_pme(S.water(nmol=64, seed=2, rigid=True), spacing=1.0, order=4)
    s = dataclasses.replace(s, thermostat="none")
    _, e, _ = orc.ewald_long_range(s, s.pos.astype(np.float64))
    r = orc.md_run(s, 1)
    r0 = orc.md_run(dataclasses.replace(s, pme=()), 1)
    assert abs(e) > 10.0
    assert abs((r["pe"][0] - r0["pe"][0]) - e) < 1e-9 * abs(e)
    with pytest.raises(ValueError):
        orc.md_run(dataclasses.replace(s, ff=dataclasses.replace(s.ff, coulomb=S.COUL_WOLF)), 1)


def test_md_with_pme_under_pressure_coupling():
    """Berendsen pressure coupling with PME: zero compressibility reproduces the
    fixed-box trajectory bit for bit (the reciprocal part follows the box of each
    step; the coupled path rebuilds its lists every step, so energies agree to the
    order of summation), a positive one moves the volume."""
    s = S.with_pme(S.nacl(nsites=12, T=900.0), spacing=1.0, order=4)
    nvt = orc.md_run(s, 20)
    frozen = orc.md_run(dataclasses.replace(s, barostat="berendsen", P0=1.0, tau_p=200.0, compress=0.0), 20)
    assert np.array_equal(frozen["pos"], nvt["pos"]) and np.array_equal(frozen["vel"], nvt["vel"])
    assert np.allclose(frozen["pe"], nvt["pe"], rtol=1e-12, atol=0.0)
    r = orc.md_run(dataclasses.replace(s, barostat="berendsen", P0=1.0, tau_p=200.0, compress=4.5e-4), 20)
    assert abs(r["volume"][-1] / np.prod(s.box.astype(np.float64)) - 1.0) > 1e-4
