"""The persistent kernel for small single systems (persist.hip): every step of
a run in one launch, the workgroups meeting at a grid barrier each step, the
lane rows built inside the kernel.

Checked against the FP64 oracle (positions, velocities, temperatures,
energies) for NVE velocity Verlet, NVE leapfrog and leapfrog Berendsen, for
each lanes-per-atom shape; against the tile path; over several calls in a row
and a hand-back to the tile path; through a row overflow (the run replayed
with longer rows); and for the 10,000-step NVE energy drift."""
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


def perturbed(sys_, amp, seed=11):
    s = sys_.copy()
    rng = np.random.default_rng(seed)
    p = s.pos.astype(np.float64) + rng.normal(0.0, amp, s.pos.shape)
    s.pos = S.wrap_centered(p, s.box.astype(np.float64)).astype(np.float32)
    return s


def arxe_one(n=2000, seed=21):
    """One Ar-Xe mixture (two atom types), NVT leapfrog Berendsen."""
    return S.arxe_replicas(nrep=1, seed=seed, nmin=n, nmax=n)[0]


CASES = [
    ("argon_vv_nve", lambda: perturbed(S.argon(), 0.1), 100),
    ("argon_lf_nve", lambda: dataclasses.replace(perturbed(S.argon(), 0.1), integrator="leapfrog",
                                                 thermostat="none"), 100),
    ("arxe2000_lf_berendsen", lambda: arxe_one(), 100),
]


def persist_engine(E, s, lanes):
    e = E.Engine.from_system(s)
    e.set_launch(persist=1, lanes=lanes)
    return e


def check_vs_oracle(st, pos, vel, r, box, pos_tol=2e-4):
    d = pos.astype(np.float64) - r["pos"]
    d -= box * np.rint(d / box)
    say("  max |dx|", np.abs(d).max(), "T", st["temp"][-1, 0], r["temp"][-1])
    assert np.abs(d).max() < pos_tol
    assert rel_rms(vel, r["vel"]) < 1e-4
    assert np.allclose(st["temp"][:, 0], r["temp"], rtol=1e-4)
    assert np.allclose(st["ke"][:, 0], r["ke"], rtol=1e-4)
    assert np.allclose(st["pe_lj"][:, 0], r["pe"], rtol=1e-4, atol=1e-4 * np.abs(r["pe"]).max())


@pytest.mark.parametrize("lanes", [4, 8, 16])
@pytest.mark.parametrize("name,mk,nsteps", CASES)
def test_persist_matches_oracle(E, orc, name, mk, nsteps, lanes):
    s = mk()
    say("persist", name, s.n, "lanes", lanes, "steps", nsteps)
    e = persist_engine(E, s, lanes)
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    st = e.run(nsteps, energy_every=1)
    info = e.persist_info()
    assert info["runs"] == 1 and info["replays"] == 0, info
    pos, vel, _ = e.download()
    r = orc.md_run(so, nsteps)
    check_vs_oracle(st, pos, vel, r, s.box)
    assert e.info()["steps"] == nsteps
    e.close()


def test_persist_matches_tile_path(E):
    s = perturbed(S.argon(), 0.1)
    say("persist vs tile path argon", s.n)
    out = []
    for persist in (0, 1):
        e = E.Engine.from_system(s)
        e.set_launch(persist=persist, lanes=16)
        st = e.run(100, energy_every=1)
        assert e.persist_info()["runs"] == persist
        pos, vel, force = e.download()
        out.append((st, pos, vel, force))
        e.close()
    (st0, p0, v0, f0), (st1, p1, v1, f1) = out
    d = p1.astype(np.float64) - p0
    d -= s.box * np.rint(d / s.box)
    say("  max |dx|", np.abs(d).max(), "force rel rms", rel_rms(f1, f0))
    assert np.abs(d).max() < 1e-4
    assert rel_rms(f1, f0) < 1e-4
    assert np.allclose(st1["conserved"][:, 0], st0["conserved"][:, 0], rtol=1e-5)


def test_persist_consecutive_runs_and_hand_back(E, orc):
    """Three persistent runs (40, 1, 39 steps), then 40 steps on the tile path:
    the 120-step trajectory of one engine against the oracle's."""
    s = dataclasses.replace(perturbed(S.argon(), 0.1), integrator="leapfrog", thermostat="berendsen", tau=100.0,
                            T0=120.0)
    say("persist consecutive runs + tile path", s.n)
    e = persist_engine(E, s, 8)
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    temps = []
    for n in (40, 1, 39):
        temps.append(e.run(n, energy_every=1)["temp"][:, 0])
    assert e.persist_info()["runs"] == 3
    e.set_launch(persist=0)
    temps.append(e.run(40, energy_every=1)["temp"][:, 0])
    assert e.persist_info()["runs"] == 3
    assert all(v == 0 for v in e.check_lists().values())
    pos, vel, _ = e.download()
    r = orc.md_run(so, 120)
    d = pos.astype(np.float64) - r["pos"]
    d -= s.box * np.rint(d / s.box)
    say("  max |dx|", np.abs(d).max())
    assert np.abs(d).max() < 2e-4
    assert np.allclose(np.concatenate(temps), r["temp"], rtol=1e-4)
    assert e.info()["steps"] == 120
    e.close()


def test_persist_row_overflow_replays(E, orc):
    """A list capacity of 16 entries per atom (about 80 neighbors lie within
    r_list) makes the lane rows too short: the run is replayed from its start with
    rows twice as long until they fit, and ends on the oracle's trajectory."""
    s = perturbed(S.argon(), 0.1)
    say("persist row overflow argon", s.n)
    e = persist_engine(E, s, 16)
    e.set_list_capacity(16)
    row0 = e.persist_info()["row_len"]
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    st = e.run(60, energy_every=1)
    info = e.persist_info()
    say("  rows per lane", row0, "->", info["row_len"], "replays", info["replays"])
    assert info["replays"] >= 1 and info["row_len"] > row0 and info["runs"] == 1
    pos, vel, _ = e.download()
    r = orc.md_run(so, 60)
    check_vs_oracle(st, pos, vel, r, s.box)
    e.close()


def test_persist_ineligible_system_runs_tiles(E, orc):
    """A charged system is not the persistent kernel's: persist = 1 runs the tile path."""
    s = dataclasses.replace(perturbed(S.nacl(nsites=12), 0.1), thermostat="none")
    say("persist ineligible nacl", s.n)
    e = persist_engine(E, s, 8)
    st = e.run(10, energy_every=1)
    assert e.persist_info()["runs"] == 0
    r = orc.md_run(s, 10)
    assert np.allclose(st["temp"][:, 0], r["temp"], rtol=1e-4)
    e.close()


def _slope_and_error(H, nblock=10):
    t = np.arange(H.size, dtype=np.float64)
    slope = np.polyfit(t, H, 1)[0]
    m = H.size // nblock
    bs = np.array([np.polyfit(t[:m], H[b * m:(b + 1) * m], 1)[0] for b in range(nblock)])
    return slope, bs.std(ddof=1) / np.sqrt(nblock)


def test_persist_nve_drift_10000_steps(E, orc):
    """Argon 864, NVE velocity Verlet dt 2 fs, 10,000 steps in one persistent
    launch: |drift| of the conserved energy <= the oracle's + 2 sigma (block
    standard errors of the two slopes)."""
    n = 10000
    s = S.argon()
    say("persist nve drift argon 10000")
    e = persist_engine(E, s, 16)
    st = e.run(n, energy_every=1)
    assert e.persist_info()["runs"] == 1
    H = st["conserved"][:, 0].astype(np.float64)
    e.close()
    r = orc.md_run(s, n)
    Ho = r["conserved"]
    sg, eg = _slope_and_error(H)
    so, eo = _slope_and_error(Ho)
    sigma = np.hypot(eg, eo)
    say("  drift per 1000 steps gpu", sg * 1e3, "+-", eg * 1e3, "oracle", so * 1e3, "+-", eo * 1e3)
    assert abs(H[0] - Ho[0]) <= 1e-4 * abs(Ho[0])
    assert abs(sg) <= abs(so) + 2.0 * sigma
