"""Parity at the benchmark sizes: the HIP engine (C ABI) against the FP64
oracle on the full configurations.

- NaCl 32,768 ions: neighbor sets bit-exact (after in-run rebuilds too);
- flexible water 196,608 atoms: forces, energies, virial, and neighbor sets
  (with the 1-2 exclusions) bit-exact;
- Ar-Xe replicas of 4,096-8,192 atoms batched in one engine: forces, sets and a
  20-step Berendsen run per replica;
- argon 864, NVE velocity Verlet, 10,000 steps: the energy drift of the GPU's
  FP32 trajectory is no larger than the FP64 oracle's plus two standard errors.
"""
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


def same_sets(counts, nbrs, oc, on, off=0):
    """Row by row set equality, vectorised: GPU rows sorted, padding past the count masked."""
    assert np.array_equal(counts, oc)
    w = max(int(oc.max()), 1)
    g = np.where(np.arange(nbrs.shape[1])[None, :] < counts[:, None], nbrs - off, np.iinfo(np.int32).max)
    g = np.sort(g, axis=1)[:, :w]
    o = np.where(np.arange(on.shape[1])[None, :] < oc[:, None], on, np.iinfo(np.int32).max)[:, :w]
    bad = np.nonzero((g != o).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first atom {bad[:5]}"


def test_neighbor_sets_nacl32768(E, orc):
    """The perturbed lattice (0.2 A) at the first build; the unperturbed lattice
    after 300 steps at 1200 K (in-run rebuilds re-sorted the atoms), with one
    more full build at the current positions so that the oracle sees the build
    positions (the perturbed lattice is only a force configuration: run, it
    explodes)."""
    say("full-size sets nacl32768")
    for stage, s, nsteps in (("perturbed, initial", perturbed(S.nacl(), 0.2), 0), ("after 300 steps", S.nacl(), 300)):
        e = E.Engine.from_system(s)
        if nsteps:
            e.run(nsteps, stats=False)
            assert e.info()["rebuilds"] > 1
            e.rebuild_bench(1)
        assert all(v == 0 for v in e.check_lists().values())
        pos, _, _ = e.download()
        _, rl2, _ = e.thresholds()
        counts, nbrs = e.neighbors()
        oc, on = orc.nlist(pos, s.box, rl2)
        say("  ", stage, "entries", int(counts.sum()))
        same_sets(counts, nbrs, oc, on)
        e.close()


def test_water196608_forces_and_sets(E, orc):
    s = perturbed(S.water(), 0.02)
    say("full-size water196608")
    e = E.Engine.from_system(s)
    pos, _, _ = e.download()
    so = s.copy()
    so.pos = pos
    F, En, W = e.compute_forces()
    Fo, Eo, Wo = orc.forces(so)
    err = rel_rms(F, Fo)
    say("  force rel rms", err, "E", En[0], Eo, "W", W[0], Wo)
    assert err < FORCE_TOL
    assert np.all(np.abs(En[0] - Eo) <= 1e-5 * np.abs(Eo).sum())
    assert abs(W[0] - Wo) <= 1e-5 * max(abs(Wo), 1.0)
    _, rl2, _ = e.thresholds()
    counts, nbrs = e.neighbors()
    oc, on = orc.nlist(pos, s.box, rl2, s.exclusions())
    say("  entries", int(counts.sum()))
    same_sets(counts, nbrs, oc, on)
    e.close()


def test_large_replicas_batched(E, orc):
    """Four Ar-Xe replicas of 4,096-8,192 atoms in one engine (the largest
    replicas of the batched benchmark)."""
    reps = S.arxe_replicas(nrep=4, seed=8, nmin=4096, nmax=8192)
    reps = [perturbed(x, 0.1, seed=30 + k) for k, x in enumerate(reps)]
    say("large replicas", [x.n for x in reps])
    e = E.Engine.from_system(reps)
    pos, _, _ = e.download()
    F, En, W = e.compute_forces()
    counts, nbrs = e.neighbors()
    off = 0
    for k, x in enumerate(reps):
        so = x.copy()
        so.pos = pos[off:off + x.n]
        Fo, Eo, Wo = orc.forces(so)
        assert rel_rms(F[off:off + x.n], Fo) < FORCE_TOL, k
        assert abs(En[k, 0] - Eo[0]) <= 1e-5 * abs(Eo[0]) + 2e-6 * x.n * EPS_MAX_ARXE
        _, rl2, _, _ = orc.thresholds(x.ff.rc, x.ff.skin, x.box)
        oc, on = orc.nlist(pos[off:off + x.n], x.box, rl2)
        same_sets(counts[off:off + x.n], nbrs[off:off + x.n], oc, on, off)
        off += x.n
    e.close()
    # 20 steps of NVT leapfrog + Berendsen from the lattice, per-replica temperatures and positions
    reps = S.arxe_replicas(nrep=4, seed=8, nmin=4096, nmax=8192)
    e = E.Engine.from_system(reps)
    st = e.run(20, energy_every=5)
    pos, vel, _ = e.download()
    off = 0
    for k, x in enumerate(reps):
        r = orc.md_run(x, 20)
        say("  replica", k, x.n, "T", st["temp"][-1, k], r["temp"][-1])
        assert np.allclose(st["temp"][:, k], r["temp"], rtol=2e-4), k
        d = pos[off:off + x.n].astype(np.float64) - r["pos"]
        d -= x.box * np.rint(d / x.box)
        assert np.abs(d).max() < 2e-4
        off += x.n
    e.close()


def _slope_and_error(H, nblock=10):
    """Least-squares slope of H(t) per step over the whole run, and its standard
    error taken as the standard error of the mean of the slopes of nblock
    consecutive blocks.  That is conservative: the whole-run fit has the smaller
    variance, and correlated fluctuations widen the block scatter as they widen
    the true error."""
    t = np.arange(H.size, dtype=np.float64)
    slope = np.polyfit(t, H, 1)[0]
    m = H.size // nblock
    bs = np.array([np.polyfit(t[:m], H[b * m:(b + 1) * m], 1)[0] for b in range(nblock)])
    return slope, bs.std(ddof=1) / np.sqrt(nblock)


def test_nve_drift_10000_steps(E, orc):
    """Argon 864, NVE velocity Verlet dt 2 fs, 10,000 steps from the same start:
    |drift| of the GPU's conserved energy <= |drift| of the oracle's + 2 sigma,
    sigma the combined standard error of the two slopes."""
    n = 10000
    s = S.argon()
    say("nve drift argon 10000")
    e = E.Engine.from_system(s)
    st = e.run(n, energy_every=1)
    H = st["conserved"][:, 0].astype(np.float64)
    e.close()
    r = orc.md_run(s, n)
    Ho = r["conserved"]
    sg, eg = _slope_and_error(H)
    so, eo = _slope_and_error(Ho)
    sigma = np.hypot(eg, eo)
    say("  drift per 1000 steps gpu", sg * 1e3, "+-", eg * 1e3, "oracle", so * 1e3, "+-", eo * 1e3,
        "fluct", np.std(Ho), "E0", H[0], Ho[0])
    assert abs(H[0] - Ho[0]) <= 1e-4 * abs(Ho[0])
    assert abs(sg) <= abs(so) + 2.0 * sigma
