"""Smooth particle-mesh Ewald on the GPU (pme.hip, through md_set_pme) against the
FP64 oracle's SPME (oracle/md_pme.cpp, pinned by the Madelung constant, beta
independence, finite differences and convergence to the direct Ewald sum).

- forces, Coulomb energy and virial for NaCl (4,096 and 32,768 ions), on cubic
  and uneven grids (odd and even edges, odd last edge), orders 4, 6 and 8, and
  in an orthorhombic box;
- rigid water (excluded intramolecular pairs) forces at 1,536 and 196,608 atoms,
  and 30 leapfrog steps with SETTLE against the oracle's trajectory;
- NaCl leapfrog Berendsen, NaCl NVE velocity Verlet (with energy conservation)
  and flexible water velocity Verlet Nose-Hoover (1-2 exclusions) trajectories;
- runs repeat bit for bit (fixed-point charge spreading), and graph-captured
  segments equal direct submission (the FFTs inside captured graphs);
- Berendsen pressure coupling with PME (the long-range part follows the box of
  each step) against the oracle's box and trajectory;
- what PME refuses: several replicas, another Coulomb kind, bad orders and
  grids."""
import dataclasses
import sys

import numpy as np
import pytest

from md_gpu_amd import systems as S
from conftest import rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def E():
    from md_gpu_amd import engine
    return engine


def perturbed(sys_, amp, seed=11):
    s = sys_.copy()
    rng = np.random.default_rng(seed)
    p = s.pos.astype(np.float64) + rng.normal(0.0, amp, s.pos.shape)
    s.pos = S.wrap_centered(p, s.box.astype(np.float64)).astype(np.float32)
    return s


def nacl_pme(nsites=16, grid=None, order=4, amp=0.2):
    s = S.with_pme(perturbed(S.nacl(nsites=nsites), amp), spacing=1.0, order=order)
    if grid is not None:
        s.pme = (*grid, order)
    return s


def water_pme(nmol=512, rigid=True, seed=2):
    return S.with_pme(S.water(nmol=nmol, seed=seed, rigid=rigid), spacing=1.0, order=4)


def check_forces(E, orc, s, tol=FORCE_TOL):
    e = E.Engine.from_system(s)
    pos, _, _ = e.download()
    so = s.copy()
    so.pos = pos
    F, En, W = e.compute_forces()
    Fo, Eo, Wo = orc.forces(so)
    err = rel_rms(F, Fo)
    say("  force rel rms", err, "E coul gpu", En[0, 1], "orc", Eo[1], "W", W[0], Wo)
    assert err < tol
    # the bar is 1e-5 of the terms summed: for neutral molecules the Coulomb energy
    # is a small remainder of the real-space, reciprocal, excluded-pair and self
    # terms, each far larger than it
    x = so.pos.astype(np.float64)
    q = so.charge.astype(np.float64)
    K0, K1, K2, order = so.pme
    e_rec = orc.pme_recip(x, q, so.box, (K0, K1, K2), order, so.ff.alpha, so.ff.qfac)[0]
    e_exc = orc.pme_excl(x, q, so.box, so.exclusions(), so.ff.alpha, so.ff.qfac)[0]
    e_self = orc.ewald_self(q, so.box, so.ff.alpha, so.ff.qfac)[0]
    e_real = Eo[1] - e_rec - e_exc - e_self
    scale = abs(Eo[0]) + abs(e_real) + abs(e_rec) + abs(e_exc) + abs(e_self)
    say("  |E gpu - E orc|", np.abs(En[0] - Eo), "terms", scale)
    assert np.all(np.abs(En[0] - Eo) <= 1e-5 * scale)
    assert abs(W[0] - Wo) <= 1e-5 * max(abs(Wo), scale)
    e.close()
    return err


@pytest.mark.parametrize("grid,order", [(None, 4), ((45, 48, 50), 4), ((48, 50, 45), 6), ((40, 40, 40), 8)])
def test_forces_nacl(E, orc, grid, order):
    s = nacl_pme(grid=grid, order=order)
    say("pme forces nacl4096", s.pme)
    check_forces(E, orc, s)


def test_forces_orthorhombic_box(E, orc):
    """Box edges 49.6, 45.1 and 42.9 A (the rock-salt lattice stretched per axis)
    on a 50 x 45 x 44 grid: every per-axis scale (grid coordinates, reciprocal
    vectors, minimum images) against the oracle."""
    s = perturbed(S.nacl(nsites=16), 0.15)
    f = np.array([1.10, 1.0, 0.95])
    s.box = (s.box.astype(np.float64) * f).astype(np.float32)
    s.pos = S.wrap_centered(s.pos.astype(np.float64) * f, s.box.astype(np.float64)).astype(np.float32)
    s = S.with_pme(s, spacing=1.0, order=4)
    s.pme = (50, 45, 44, 4)
    say("pme forces orthorhombic", s.box, s.pme)
    check_forces(E, orc, s)


def test_forces_nacl_full_size(E, orc):
    s = nacl_pme(nsites=32)
    say("pme forces nacl32768", s.pme)
    check_forces(E, orc, s)


def test_forces_rigid_water_full_size(E, orc):
    """65,536 rigid molecules (196,608 atoms) on a 128^3 grid: the bench workload's
    forces, energy and virial against the oracle."""
    s = water_pme(nmol=65536, seed=3)
    say("pme forces rigid water full size", s.n, s.pme)
    check_forces(E, orc, s)


def test_long_range_part_is_large(E, orc):
    """The reciprocal part changes the forces by far more than the tolerance: the
    force test is no test of the real-space kernel alone."""
    s = nacl_pme()
    e = E.Engine.from_system(s)
    F1, _, _ = e.compute_forces()
    e.set_pme((1, 1, 1), 0)
    F0, _, _ = e.compute_forces()
    d = rel_rms(F1, F0)
    say("  |F_pme - F_real| rel", d)
    assert d > 1e-3
    e.close()


@pytest.mark.parametrize("order", [4, 6])
def test_forces_rigid_water(E, orc, order):
    s = water_pme()
    s.pme = (*s.pme[:3], order)
    say("pme forces rigid water", s.n, s.pme)
    check_forces(E, orc, s)


@pytest.mark.parametrize("alpha", [0.2, 0.35])
def test_forces_nacl_splitting_parameter(E, orc, alpha):
    """Other real/reciprocal splits than alpha = 3.09 / rc: the damped table, the
    influence function and the self term all follow alpha."""
    s = S.with_pme(perturbed(S.nacl(nsites=16), 0.2), spacing=0.8 if alpha > 0.3 else 1.0, order=4, alpha=alpha)
    say("pme forces nacl alpha", alpha, s.pme)
    check_forces(E, orc, s)


def equilibrated_rigid_water(orc, nmol=512, seed=2, steps=150):
    s = S.water(nmol=nmol, seed=seed, rigid=True)
    r = orc.md_run(dataclasses.replace(s, dt=1.0, tau=20.0), steps)
    t = s.copy()
    t.pos = r["pos"].astype(np.float32)
    t.vel = r["vel"].astype(np.float32)
    return S.with_pme(t, spacing=1.0, order=4)


def run_vs_oracle(E, orc, s, n, launch=None, tol_x=2e-4):
    e = E.Engine.from_system(s)
    if launch:
        e.set_launch(**launch)
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    st = e.run(n, energy_every=1)
    pos, vel, _ = e.download()
    r = orc.md_run(so, n)
    d = pos.astype(np.float64) - r["pos"]
    d -= s.box * np.rint(d / s.box)
    say("  max |dx|", np.abs(d).max(), "T", st["temp"][-1, 0], r["temp"][-1])
    assert np.abs(d).max() < tol_x
    assert rel_rms(vel, r["vel"]) < 2e-4
    assert np.allclose(st["temp"][:, 0], r["temp"], rtol=2e-4)
    pe = st["pe_lj"][:, 0] + st["pe_coul"][:, 0] + st["pe_bond"][:, 0] + st["pe_angle"][:, 0]
    assert np.allclose(pe, r["pe"], rtol=1e-4, atol=1e-4 * np.abs(r["pe"]).max())
    e.close()


@pytest.mark.parametrize("launch", [dict(lanes=8), dict(lanes=4, submit=1, check_every=4)])
def test_rigid_water_trajectory(E, orc, launch):
    s = equilibrated_rigid_water(orc)
    say("pme rigid water 30 steps", s.n, launch)
    run_vs_oracle(E, orc, s, 30, launch)


def test_nacl_berendsen_trajectory(E, orc):
    s = nacl_pme(amp=0.1)
    say("pme nacl leapfrog berendsen 20 steps")
    run_vs_oracle(E, orc, s, 20)


def test_flexible_water_vv_nh_trajectory(E, orc):
    s = S.with_pme(perturbed(S.water(nmol=512), 0.02), spacing=1.0, order=4)
    say("pme flexible water vv nh 20 steps")
    run_vs_oracle(E, orc, s, 20)


def test_nacl_nve_vv_trajectory_and_conservation(E, orc):
    """NVE velocity Verlet runs in leapfrog form (one step kernel per step): 20 steps
    against the oracle, then 400 more conserving the total energy within the
    oracle's own bounds (drift < 0.2 %, span < 0.5 % of the kinetic energy)."""
    s = dataclasses.replace(S.with_pme(S.nacl(nsites=12, T=600.0), spacing=1.0, order=4),
                            thermostat="none", integrator="vverlet", dt=1.0)
    say("pme nacl nve vv 20 + 400 steps")
    run_vs_oracle(E, orc, s, 20)
    e = E.Engine.from_system(s)
    n = 400
    st = e.run(n, energy_every=1)
    H = st["conserved"][:, 0]
    ke = st["ke"][:, 0].mean()
    slope = np.polyfit(np.arange(n), H, 1)[0]
    say("  drift", slope * n, "span", H.max() - H.min(), "ke", ke)
    assert abs(slope) * n < 0.002 * ke
    assert H.max() - H.min() < 0.005 * ke
    e.close()


def test_npt_nacl_with_pme(E, orc):
    """Leapfrog, Berendsen temperature and pressure coupling, 40 steps: the box,
    the volume series, the temperature and the positions follow the oracle."""
    s = dataclasses.replace(nacl_pme(nsites=12, amp=0.1), barostat="berendsen", P0=1.0, tau_p=200.0,
                            compress=4.5e-4)
    n = 40
    say("pme npt nacl", s.n, s.pme)
    e = E.Engine.from_system(s)
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    st = e.run(n, energy_every=1)
    pos, _, _ = e.download()
    box = e.box()[0].astype(np.float64)
    r = orc.md_run(so, n)
    say("  box gpu", box, "oracle", r["box"], "V0", np.prod(s.box))
    assert abs(np.prod(r["box"]) / np.prod(s.box.astype(np.float64)) - 1.0) > 1e-4   # the box did move
    np.testing.assert_allclose(box, r["box"], rtol=2e-5)
    np.testing.assert_allclose(st["volume"][:, 0], r["volume"], rtol=2e-5)
    np.testing.assert_allclose(st["temp"][:, 0], r["temp"], rtol=2e-4)
    d = pos.astype(np.float64) - r["pos"]
    d -= box * np.rint(d / box)
    assert np.abs(d).max() < 5e-4
    assert all(v == 0 for v in e.check_lists().values())
    e.close()


def test_runs_repeat_and_graphs_equal_direct(E):
    """Fixed-point spreading makes the charge grid independent of the order of the
    atomic adds: two runs from one state agree bit for bit, and so do graph
    segments (FFTs captured and replayed, several segments) and direct submission."""
    s = nacl_pme(amp=0.1)
    out = []
    for launch in (dict(submit=0, check_every=5), dict(submit=0, check_every=5), dict(submit=1, check_every=5)):
        e = E.Engine.from_system(s)
        e.set_launch(**launch)
        for _ in range(3):
            e.run(20, energy_every=5)
        pos, vel, _ = e.download()
        out.append((pos, vel))
        e.close()
    for pos, vel in out[1:]:
        assert np.array_equal(pos, out[0][0]) and np.array_equal(vel, out[0][1])


def test_refusals(E):
    s = nacl_pme(nsites=12, amp=0.05)
    # several replicas
    with pytest.raises(E.MDError):
        E.Engine.from_system([s, s.copy()])
    # another Coulomb kind
    t = dataclasses.replace(s.copy(), ff=dataclasses.replace(s.ff, coulomb=S.COUL_WOLF))
    e = E.Engine.from_system(t)
    with pytest.raises(E.MDError):
        e.run(1)
    e.close()
    # bad orders and grids
    e = E.Engine.from_system(s)
    for grid, order in (((32, 32, 32), 5), ((3, 32, 32), 4), ((32, 32, 32), 10)):
        with pytest.raises(E.MDError):
            e.set_pme(grid, order)
    e.close()
