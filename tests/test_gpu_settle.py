"""Rigid 3-site water on the GPU (SETTLE after every leapfrog step, settle.hip)
against the FP64 oracle's SETTLE, which the CPU tests pin to SHAKE, exact
geometry, the centre of mass and NVE conservation.

- 512 molecules from an equilibrated state: positions, velocities,
  temperatures (3N - 3 - 3 N_mol degrees of freedom) and energies over 30 steps,
  for several launch shapes (deferred finalize is set aside for rigid systems);
- two rigid replicas of different sizes in one engine;
- the full 65,536 molecules (196,608 atoms): 10 steps against the oracle from
  the lattice start, and the geometry after 200 steps;
- runs with velocity Verlet are refused."""
import dataclasses
import sys

import numpy as np
import pytest

from md_gpu_amd import systems as S
from conftest import rel_rms

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def E():
    from md_gpu_amd import engine
    return engine


def dists(x, mol, box):
    L = np.asarray(box, np.float64)
    o, h1, h2 = mol.T

    def d(i, j):
        v = x[j].astype(np.float64) - x[i]
        v -= L * np.rint(v / L)
        return np.linalg.norm(v, axis=1)

    return d(o, h1), d(o, h2), d(h1, h2)


def geometry_error(x, s):
    return max(np.abs(g - w).max() for g, w in zip(dists(x, s.rigid, s.box), (s.rigid_d[0], s.rigid_d[0], s.rigid_d[1])))


def equilibrated(orc, nmol=512, seed=2, steps=200):
    """Rigid water after `steps` oracle steps of Berendsen at 300 K (tau 20 fs,
    dt 1 fs) from the lattice start, as float32 state."""
    s = S.water(nmol=nmol, seed=seed, rigid=True)
    r = orc.md_run(dataclasses.replace(s, dt=1.0, tau=20.0), steps)
    t = s.copy()
    t.pos = r["pos"].astype(np.float32)
    t.vel = r["vel"].astype(np.float32)
    return t


@pytest.mark.parametrize("launch", [dict(lanes=8), dict(lanes=4, submit=1, check_every=4),
                                    dict(lanes=16, defer=1, refresh=1, check_every=6)])
def test_rigid_water_matches_oracle(E, orc, launch):
    s = equilibrated(orc)
    n = 30
    say("rigid water", s.n, launch)
    e = E.Engine.from_system(s)
    e.set_launch(**launch)
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    st = e.run(n, energy_every=1)
    pos, vel, _ = e.download()
    r = orc.md_run(so, n)
    d = pos.astype(np.float64) - r["pos"]
    d -= s.box * np.rint(d / s.box)
    say("  max |dx|", np.abs(d).max(), "geometry", geometry_error(pos, s), "T", st["temp"][-1, 0], r["temp"][-1])
    assert np.abs(d).max() < 2e-4
    assert rel_rms(vel, r["vel"]) < 2e-4
    assert np.allclose(st["temp"][:, 0], r["temp"], rtol=2e-4)
    pe = st["pe_lj"][:, 0] + st["pe_coul"][:, 0]
    assert np.allclose(pe, r["pe"], rtol=1e-4, atol=1e-4 * np.abs(r["pe"]).max())
    assert geometry_error(pos, s) < 2e-5
    assert all(v == 0 for v in e.check_lists().values())
    e.close()


def test_rigid_water_replicas(E, orc):
    reps = [equilibrated(orc, nmol=512, seed=2, steps=100), equilibrated(orc, nmol=729, seed=4, steps=100)]
    n = 20
    say("rigid replicas", [x.n for x in reps])
    e = E.Engine.from_system(reps)
    pos0, vel0, _ = e.download()
    st = e.run(n, energy_every=1)
    pos, _, _ = e.download()
    off = 0
    for k, x in enumerate(reps):
        so = x.copy()
        so.pos, so.vel = pos0[off:off + x.n], vel0[off:off + x.n]
        r = orc.md_run(so, n)
        d = pos[off:off + x.n].astype(np.float64) - r["pos"]
        d -= x.box * np.rint(d / x.box)
        say("  replica", k, "max |dx|", np.abs(d).max(), "T", st["temp"][-1, k], r["temp"][-1])
        assert np.abs(d).max() < 2e-4
        assert np.allclose(st["temp"][:, k], r["temp"], rtol=2e-4)
        assert geometry_error(pos[off:off + x.n], x) < 2e-5
        off += x.n
    e.close()


def test_rigid_water_full_size(E, orc):
    """65,536 rigid molecules (196,608 atoms) from the lattice start, dt 1 fs,
    Berendsen tau 20 fs: 10 steps against the oracle, then 190 more with the
    geometry held and the temperature settling near 300 K."""
    s = dataclasses.replace(S.water(rigid=True), dt=1.0, tau=20.0)
    say("rigid water full size", s.n)
    e = E.Engine.from_system(s)
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    st = e.run(10, energy_every=1)
    pos, _, _ = e.download()
    r = orc.md_run(so, 10)
    d = pos.astype(np.float64) - r["pos"]
    d -= s.box * np.rint(d / s.box)
    say("  10 steps: max |dx|", np.abs(d).max(), "T", st["temp"][-1, 0], r["temp"][-1])
    assert np.abs(d).max() < 5e-4
    assert np.allclose(st["temp"][:, 0], r["temp"], rtol=5e-4)
    st = e.run(190, energy_every=10)
    pos, _, _ = e.download()
    g = geometry_error(pos, s)
    say("  200 steps: geometry", g, "T", st["temp"][-1, 0], "rebuilds", e.info()["rebuilds"])
    assert g < 2e-5
    assert 250.0 < st["temp"][-1, 0] < 450.0
    e.close()


def test_rigid_needs_leapfrog(E):
    s = dataclasses.replace(S.water(nmol=512, rigid=True), integrator="vverlet", thermostat="none")
    e = E.Engine.from_system(s)
    with pytest.raises(E.MDError):
        e.run(2)
    e.close()
