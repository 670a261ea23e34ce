"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle.

Bars (north star): neighbor sets bit-exact; per-atom forces within 1e-5
relative RMS; energies and virial within 1e-5 relative; integrator
trajectories within FP32 round-off of the FP64 oracle over short runs.
"""
import dataclasses
import sys

import numpy as np
import pytest

from md_gpu_amd import systems as S
from conftest import rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
EPS_MAX_ARXE = 221.0 * S.KB_KJ


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


def perturbed(sys_, amp, seed=11):
    s = sys_.copy()
    rng = np.random.default_rng(seed)
    p = s.pos.astype(np.float64) + rng.normal(0.0, amp, s.pos.shape)
    s.pos = S.wrap_centered(p, s.box.astype(np.float64)).astype(np.float32)
    return s


def water_geometry_perturbed(sys_, amp, seed=12):
    return perturbed(sys_, amp, seed)


def make_cases():
    ar = perturbed(S.argon(), 0.25)
    na = perturbed(S.nacl(nsites=16), 0.2)
    na_rf = dataclasses.replace(na.copy(), ff=dataclasses.replace(S.nacl_ff(), coulomb=S.COUL_RF, eps_rf=78.0))
    na_d = dataclasses.replace(na.copy(), ff=dataclasses.replace(S.nacl_ff(), coulomb=S.COUL_DAMPED, alpha=0.25))
    wa = perturbed(S.water(nmol=512), 0.05)
    return {"argon864": ar, "nacl4096_wolf": na, "nacl4096_rf": na_rf, "nacl4096_damped": na_d, "water1536": wa}


CASES = make_cases()


@pytest.fixture(scope="module")
def E():
    from md_gpu_amd import engine
    return engine


def oracle_forces(orc, s):
    return orc.forces(s)


@pytest.mark.parametrize("lanes,cell_div,build", [(4, 1, 4), (8, 2, 8), (16, 3, 4)])
@pytest.mark.parametrize("name", list(CASES))
def test_neighbor_sets_bit_exact(E, orc, name, lanes, cell_div, build):
    s = CASES[name]
    say("neighbors", name, s.n, "lanes", lanes, "cell_div", cell_div, "build wavefronts", build)
    e = E.Engine.from_system(s)
    e.set_launch(lanes=lanes, cell_div=cell_div, build_lanes=build)
    assert all(v == 0 for v in e.check_lists().values()), e.check_lists()
    pos, _, _ = e.download()
    rc2, rl2, sk = e.thresholds()
    orc_rc2, orc_rl2, orc_sk, _ = orc.thresholds(s.ff.rc, s.ff.skin, s.box)
    assert (rc2, rl2, sk) == (orc_rc2, orc_rl2, orc_sk)
    counts, nbrs = e.neighbors()
    oc, on = orc.nlist(pos, s.box, rl2, s.exclusions())
    assert np.array_equal(counts, oc)
    for i in range(s.n):
        got = np.sort(nbrs[i, :counts[i]])
        assert np.array_equal(got, on[i, :oc[i]]), f"atom {i}"
    e.close()


@pytest.mark.parametrize("name", ["argon864", "nacl4096_wolf", "water1536"])
def test_build_wavefronts_same_lists(E, name):
    """The tile build with 4 or 8 wavefronts per tile writes the same sets and rows
    (candidate order, not scheduling, fixes them): identical neighbor arrays, and
    identical 200-step trajectories through in-run rebuilds."""
    s = CASES[name]
    say("build wavefronts", name, s.n)
    out = []
    for nw in (4, 8):
        e = E.Engine.from_system(s)
        e.set_launch(build_lanes=nw, persist=0)
        counts, nbrs = e.neighbors()
        st = e.run(200, energy_every=10)
        pos, vel, _ = e.download()
        out.append((counts, nbrs, st, pos, vel, e.info()["rebuilds"]))
        e.close()
    a, b = out
    assert a[5] == b[5] and a[5] > 1
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for nm in a[2].dtype.names:
        assert np.array_equal(a[2][nm], b[2][nm]), nm
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])


@pytest.mark.parametrize("name", list(CASES))
def test_forces_energy_virial(E, orc, name):
    s = CASES[name]
    say("forces", name, s.n)
    e = E.Engine.from_system(s)
    pos, _, _ = e.download()
    so = s.copy()
    so.pos = pos
    F, En, W = e.compute_forces()
    Fo, Eo, Wo = orc.forces(so)
    err = rel_rms(F, Fo)
    say("  force rel rms", err, "E gpu", En[0], "E orc", Eo, "W", W[0], Wo)
    assert err < FORCE_TOL
    scale = max(np.abs(Eo).sum(), 1.0)
    assert np.all(np.abs(En[0] - Eo) <= 1e-5 * scale)
    assert abs(W[0] - Wo) <= 1e-5 * max(abs(Wo), np.abs(Fo).sum() * 1e-3, 1.0)
    # momentum: sum of forces vanishes (full lists, Newton III in aggregate)
    assert np.abs(F.astype(np.float64).sum(0)).max() <= 1e-4 * np.abs(F).sum(0).max() + 1e-3
    e.close()


def test_forces_full_size_nacl(E, orc):
    s = perturbed(S.nacl(), 0.2)
    say("forces nacl32768")
    e = E.Engine.from_system(s)
    pos, _, _ = e.download()
    so = s.copy()
    so.pos = pos
    F, En, W = e.compute_forces()
    Fo, Eo, Wo = orc.forces(so)
    err = rel_rms(F, Fo)
    say("  force rel rms", err)
    assert err < FORCE_TOL
    e.close()


INTEG_CASES = [
    ("argon_vv_nve", lambda: perturbed(S.argon(), 0.1), 20),
    ("nacl_lf_berendsen", lambda: perturbed(S.nacl(nsites=16), 0.1), 20),
    ("nacl_lf_nve", lambda: dataclasses.replace(perturbed(S.nacl(nsites=16), 0.1), thermostat="none"), 20),
    ("water_vv_nh", lambda: perturbed(S.water(nmol=512), 0.02), 20),
    ("nacl_vv_berendsen", lambda: dataclasses.replace(perturbed(S.nacl(nsites=16), 0.1), integrator="vverlet"), 20),
]


@pytest.mark.parametrize("name,mk,nsteps", INTEG_CASES)
def test_integrator_short_run(E, orc, name, mk, nsteps):
    s = mk()
    say("integrate", name, s.n, nsteps)
    e = E.Engine.from_system(s)
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    st = e.run(nsteps, energy_every=1)
    pos, vel, _ = e.download()
    r = orc.md_run(so, nsteps)
    dpos = pos.astype(np.float64) - r["pos"]
    dpos -= s.box * np.rint(dpos / s.box)
    say("  max |dx|", np.abs(dpos).max(), "T gpu", st["temp"][-1, 0], "T orc", r["temp"][-1])
    assert np.abs(dpos).max() < 2e-4
    assert rel_rms(vel, r["vel"]) < 1e-4
    assert np.allclose(st["temp"][:, 0], r["temp"], rtol=1e-4)
    assert np.allclose(st["ke"][:, 0], r["ke"], rtol=1e-4)
    pe = st["pe_lj"][:, 0] + st["pe_coul"][:, 0] + st["pe_bond"][:, 0] + st["pe_angle"][:, 0]
    assert np.allclose(pe, r["pe"], rtol=1e-4, atol=1e-4 * np.abs(r["pe"]).max())
    e.close()


def test_nve_energy_drift(E, orc):
    """Argon, NVE velocity Verlet dt 2 fs, 1000 steps: a loose sanity bound (the GPU
    drift within twice the oracle's, half its fluctuation, or 1e-6 of the energy).
    Over 1000 steps the drift is the size of the fluctuation, so this is no
    statistical comparison; test_gpu_fullsize.py::test_nve_drift_10000_steps (and
    the persistent kernel's test_persist_nve_drift_10000_steps) make that one:
    |drift_gpu| <= |drift_oracle| + 2 sigma over 10,000 steps."""
    s = S.argon()
    say("nve drift argon 1000")
    e = E.Engine.from_system(s)
    st = e.run(1000, energy_every=1)
    H = st["conserved"][:, 0]
    r = orc.md_run(s, 1000)
    Ho = r["conserved"]
    t = np.arange(1000)
    drift_g = abs(np.polyfit(t, H, 1)[0]) * 1000
    drift_o = abs(np.polyfit(t, Ho, 1)[0]) * 1000
    fluct_o = np.std(Ho)
    say("  drift gpu", drift_g, "drift orc", drift_o, "fluct", fluct_o, "E0", H[0], Ho[0])
    assert drift_g <= max(2.0 * drift_o, 0.5 * fluct_o, 1e-6 * abs(Ho[0]))
    e.close()


def test_batched_replicas_forces(E, orc):
    reps = [perturbed(x, 0.1, seed=20 + k) for k, x in enumerate(S.arxe_replicas(nrep=12, seed=5, nmax=2048))]
    say("batch forces", [x.n for x in reps])
    e = E.Engine.from_system(reps)
    pos, _, _ = e.download()
    F, En, W = e.compute_forces()
    off = 0
    for k, x in enumerate(reps):
        so = x.copy()
        so.pos = pos[off:off + x.n]
        Fo, Eo, Wo = orc.forces(so)
        assert rel_rms(F[off:off + x.n], Fo) < FORCE_TOL, k
        # FP32 pair energies summed in FP64: bound per atom in units of the largest epsilon
        assert abs(En[k, 0] - Eo[0]) <= 1e-5 * abs(Eo[0]) + 2e-6 * x.n * EPS_MAX_ARXE
        off += x.n
    counts, nbrs = e.neighbors()
    off = 0
    for k, x in enumerate(reps):
        rc2, rl2, _, _ = orc.thresholds(x.ff.rc, x.ff.skin, x.box)
        oc, on = orc.nlist(pos[off:off + x.n], x.box, rl2)
        assert np.array_equal(counts[off:off + x.n], oc)
        for i in range(x.n):
            assert np.array_equal(np.sort(nbrs[off + i, :counts[off + i]]) - off, on[i, :oc[i]])
        off += x.n
    e.close()


def test_batched_replicas_run(E, orc):
    """Per-replica Berendsen temperatures and trajectories match independent oracle runs."""
    reps = S.arxe_replicas(nrep=6, seed=6, nmax=1024)
    say("batch run", [x.n for x in reps])
    e = E.Engine.from_system(reps)
    st = e.run(20, energy_every=5)
    pos, vel, _ = e.download()
    off = 0
    for k, x in enumerate(reps):
        r = orc.md_run(x, 20)
        assert np.allclose(st["temp"][:, k], r["temp"], rtol=2e-4), k
        d = pos[off:off + x.n].astype(np.float64) - r["pos"]
        d -= x.box * np.rint(d / x.box)
        assert np.abs(d).max() < 2e-4
        off += x.n
    e.close()


@pytest.mark.parametrize("fused", [1, 0])
def test_force_matching_gradient(E, orc, fused):
    """Loss and gradient of the force-matching loss against the FP64 oracle,
    for the single pass (force and gradient sums together) and for the force
    pass followed by the adjoint walk."""
    base = S.nacl(nsites=16)
    rng = np.random.default_rng(3)
    nf = 4
    frames = np.stack([S.wrap_centered(base.pos + rng.normal(0, 0.25, base.pos.shape), base.box)
                       for _ in range(nf)]).astype(np.float32)
    tgt_ff, tgt_q = S.perturb_nacl_ff(base.ff, seed=9)
    targets = []
    for f in range(nf):
        s = base.copy()
        s.pos = frames[f]
        s.charge = base.charge * tgt_q
        Ft, _, _ = orc.forces(dataclasses.replace(s, ff=tgt_ff))
        targets.append(Ft)
    targets = np.stack(targets)
    say("fm", nf, base.n, "fused", fused)
    e = E.Engine.from_system(base)
    e.set_launch(fm_fused=fused)
    e.fm_set_frames(frames, targets)
    loss, g1, g2, gq = e.fm_loss_grad(base.ff.lj1, base.ff.lj2, 1.0)
    lo, o1, o2, oq = orc.fm_loss_grad(frames, base.type, base.charge, base.box, base.ff, 1.0, targets)
    say("  loss", loss, lo, "gq", gq, oq)
    assert abs(loss - lo) <= 1e-4 * lo
    assert abs(gq - oq) <= 1e-4 * abs(oq)
    assert np.allclose(g1, o1, rtol=1e-4, atol=1e-4 * np.abs(o1).max())
    assert np.allclose(g2, o2, rtol=1e-4, atol=1e-4 * np.abs(o2).max())
    e.close()


@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_force_matching_single_pass_matches_two_pass(E, lanes):
    """The single-pass loss and gradient equal the force pass + adjoint walk to
    FP32 summation-order differences, at the base and at perturbed parameters."""
    base = S.nacl(nsites=16)
    rng = np.random.default_rng(5)
    nf = 3
    frames = np.stack([S.wrap_centered(base.pos + rng.normal(0, 0.25, base.pos.shape), base.box)
                       for _ in range(nf)]).astype(np.float32)
    targets = rng.normal(0, 50.0, (nf, base.n, 3)).astype(np.float32)
    tgt_ff, tgt_q = S.perturb_nacl_ff(base.ff, seed=11)
    say("fm single vs two pass", lanes)
    out = {}
    for fused in (1, 0):
        e = E.Engine.from_system(base)
        e.set_launch(fm_fused=fused, lanes=lanes)
        e.fm_set_frames(frames, targets)
        out[fused] = [e.fm_loss_grad(base.ff.lj1, base.ff.lj2, 1.0), e.fm_loss_grad(tgt_ff.lj1, tgt_ff.lj2, tgt_q)]
        e.close()
    for a, b in zip(out[1], out[0]):
        la, g1a, g2a, gqa = a
        lb, g1b, g2b, gqb = b
        say("  loss", la, lb, "gq", gqa, gqb)
        assert abs(la - lb) <= 1e-6 * lb
        assert abs(gqa - gqb) <= 1e-5 * abs(gqb)
        assert np.allclose(g1a, g1b, rtol=1e-5, atol=1e-6 * np.abs(g1b).max())
        assert np.allclose(g2a, g2b, rtol=1e-5, atol=1e-6 * np.abs(g2b).max())


def test_logged_reference_forces_nacl_rf(E, orc):
    """The engine reproduces the step-0 forces the reference logged for
    examples/nacl (reaction field, 4096 ions), rows i % 128 == 0."""
    from test_oracle_golden import load_fixture
    s, _, idx, F_ref = load_fixture()
    say("logged forces nacl_rf_step0", s.n)
    e = E.Engine.from_system(s)
    F, _, _ = e.compute_forces()
    rel = np.linalg.norm(F[idx] - F_ref, axis=1) / np.linalg.norm(F_ref, axis=1)
    Fo, _, _ = orc.forces(s)
    say("  max rel vs log", rel.max(), "rel rms vs oracle", rel_rms(F, Fo))
    assert rel.max() < 3e-5
    assert rel_rms(F, Fo) < FORCE_TOL
    e.close()


LAUNCH_VARIANTS = [
    dict(submit=1, refresh=0, check_every=4, lanes=4),
    dict(submit=0, refresh=1, check_every=4, lanes=8),
    dict(submit=1, refresh=1, check_every=6, lanes=16),
    dict(submit=0, refresh=0, check_every=12, lanes=16, cell_div=3),
    dict(submit=0, refresh=0, check_every=8, lanes=8, cell_div=1),
    dict(submit=0, refresh=1, check_every=8, lanes=4, cell_div=2, defer=1),
    dict(submit=1, refresh=0, check_every=16, lanes=16, cell_div=3, defer=0),
    dict(submit=0, refresh=1, check_every=6, lanes=8, cell_div=1, defer=1),
    dict(submit=1, refresh=0, check_every=12, lanes=4, defer=1),
]


@pytest.mark.parametrize("variant", range(len(LAUNCH_VARIANTS)))
@pytest.mark.parametrize("kind", ["nacl_lf_berendsen", "water_vv_nh"])
def test_launch_variants_match_oracle(E, orc, kind, variant):
    """Submission mode, age refresh and kernel shapes change only the schedule,
    never the result: every variant follows the oracle trajectory."""
    cfg = LAUNCH_VARIANTS[variant]
    s = perturbed(S.nacl(nsites=16), 0.1) if kind == "nacl_lf_berendsen" else perturbed(S.water(nmol=512), 0.02)
    say("launch variant", kind, cfg)
    e = E.Engine.from_system(s)
    e.set_launch(**cfg)
    assert {k: e.get_launch()[k] for k in cfg} == cfg
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    nsteps = 30
    st = e.run(nsteps, energy_every=10)
    pos, vel, _ = e.download()
    r = orc.md_run(so, nsteps)
    dpos = pos.astype(np.float64) - r["pos"]
    dpos -= s.box * np.rint(dpos / s.box)
    say("  max |dx|", np.abs(dpos).max(), "rebuilds", e.info()["rebuilds"])
    assert np.abs(dpos).max() < 2e-4
    assert np.allclose(st["temp"][:, 0], r["temp"], rtol=1e-4)
    e.close()
