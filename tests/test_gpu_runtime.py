"""GPU checks of state that only exists during runs: neighbor lists rebuilt
inside replayed graph segments, run-to-run reproducibility, and energy
conservation at the benchmark size."""
import dataclasses
import sys

import numpy as np
import pytest

from md_gpu_amd import systems as S

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def E():
    from md_gpu_amd import engine
    return engine


def pairs_within(orc, pos, box, r2):
    counts, nb = orc.nlist(pos, box, float(np.float32(r2)))
    return [set(nb[i, :counts[i]].tolist()) for i in range(len(pos))]


@pytest.mark.parametrize("nsites,cfg", [(16, dict(check_every=8)),
                                        (32, dict(check_every=16, lanes=8)),
                                        (32, dict(check_every=16, lanes=4, cell_div=1)),
                                        (32, dict(check_every=8, lanes=8, cell_div=3)),
                                        (32, dict(check_every=8, lanes=16)),
                                        (32, dict(check_every=4, refresh=1, submit=1))])
def test_lists_valid_after_in_run_rebuilds(E, orc, nsites, cfg):
    """After a run with many device-side rebuilds, the device's sorted order and
    lists pass the invariant check (order a bijection, every listed pair within
    r_list at the build positions, sets and padding sound), every pair within rc
    of the current positions is in the list, and no row holds a pair beyond
    rc + 2 skin (the list is never older than a skin/2 move per atom)."""
    s = S.nacl(nsites=nsites)
    say("lists after run", s.n, cfg)
    e = E.Engine.from_system(s)
    e.set_launch(**cfg)
    reb0 = e.info()["rebuilds"]
    for _ in range(3):
        e.run(40, energy_every=0, stats=False)
        bad = e.check_lists()
        assert all(v == 0 for v in bad.values()), bad
    assert e.info()["rebuilds"] - reb0 >= 3
    pos, _, _ = e.download()
    counts, nbrs = e.neighbors()
    rc2, _, _, _ = orc.thresholds(s.ff.rc, s.ff.skin, s.box)
    need = pairs_within(orc, pos, s.box, rc2)
    outer = pairs_within(orc, pos, s.box, (s.ff.rc + 2 * s.ff.skin) ** 2)
    for i in range(s.n):
        row = set(nbrs[i, :counts[i]].tolist())
        assert need[i] <= row, f"atom {i}: missing {sorted(need[i] - row)[:5]}"
        assert row <= outer[i], f"atom {i}: stale {sorted(row - outer[i])[:5]}"
    e.close()


def test_runs_reproducible(E):
    """Two engines on the same input follow the same trajectory bit for bit:
    every reduction has a fixed order (no floating-point atomics) and the cell
    sort orders each cell by original index."""
    s = S.nacl()
    say("reproducible nacl32768")
    runs = []
    for _ in range(2):
        e = E.Engine.from_system(s)
        e.set_launch(lanes=8, check_every=16)
        st = e.run(150, energy_every=10)
        pos, vel, _ = e.download()
        runs.append((st["temp"][:, 0].copy(), pos, e.info()["rebuilds"]))
        e.close()
    (t0, p0, r0), (t1, p1, r1) = runs
    say("  rebuilds", r0, r1, "max dT/T", np.abs(t0 - t1).max() / t0.mean(), "max dx", np.abs(p0 - p1).max())
    assert r0 == r1
    assert np.array_equal(t0, t1)
    assert np.array_equal(p0, p1)


@pytest.mark.parametrize("cfg", [dict(check_every=8, refresh=1), dict(check_every=6, refresh=0, submit=1),
                                 dict(check_every=16, refresh=0, lanes=4)])
def test_deferred_finalize_bit_identical(E, cfg):
    """Leapfrog blocks whose step kernels finalize the previous step (defer=1)
    follow the same trajectory bit for bit as a finalize kernel after every step:
    positions, velocities, forces and every statistics record, through list
    rebuilds, stalls on a stale list and energy steps."""
    s = S.nacl(nsites=16)
    say("deferred finalize", s.n, cfg)
    out = []
    for defer in (0, 1):
        e = E.Engine.from_system(s)
        e.set_launch(defer=defer, **cfg)
        st = e.run(203, energy_every=7)
        st2 = e.run(50, energy_every=0)
        pos, vel, frc = e.download()
        out.append((st, st2, pos, vel, frc, e.info()["rebuilds"]))
        e.close()
    a, b = out
    say("  rebuilds", a[5], b[5])
    assert a[5] == b[5] and a[5] > 10
    for k in range(5):
        if k < 2:
            for name in a[k].dtype.names:
                assert np.array_equal(a[k][name], b[k][name]), name
        else:
            assert np.array_equal(a[k], b[k])


@pytest.mark.parametrize("lanes", [4, 16])
def test_deferred_finalize_bit_identical_batched(E, lanes):
    """The same for six Ar-Xe replicas of uneven size: every workgroup sums the
    slots of its replica's tiles."""
    reps = S.arxe_replicas(nrep=6, seed=6, nmax=1024)
    say("deferred finalize batched", [x.n for x in reps], "lanes", lanes)
    out = []
    for defer in (0, 1):
        e = E.Engine.from_system(reps)
        e.set_launch(defer=defer, lanes=lanes, check_every=8, refresh=1)
        st = e.run(120, energy_every=5)
        pos, vel, frc = e.download()
        out.append((st, pos, vel, frc, e.info()["rebuilds"]))
        e.close()
    a, b = out
    say("  rebuilds", a[4], b[4])
    assert a[4] == b[4] and a[4] > 6
    for name in a[0].dtype.names:
        assert np.array_equal(a[0][name], b[0][name]), name
    for k in range(1, 4):
        assert np.array_equal(a[k], b[k])


def _nve_vv(s):
    s = s.copy()
    s.integrator, s.thermostat = "vverlet", "none"
    return s


@pytest.mark.parametrize("name,cfg", [("nacl", dict(check_every=8)),
                                      ("nacl", dict(check_every=6, defer=0, submit=1)),
                                      ("nacl", dict(check_every=4, lanes=16)),
                                      ("argon", dict(check_every=16, lanes=16))])
def test_fused_vv_bit_identical(E, name, cfg):
    """NVE velocity Verlet in leapfrog form (one step kernel per step that ends
    the step and begins the next, vv_fuse=1) follows the same trajectory bit for
    bit as the kick / force / finalize kernels (vv_fuse=0): positions, full-step
    velocities, forces and every statistics record, through list rebuilds,
    stalls on a stale list, energy steps and two runs."""
    s = _nve_vv(S.nacl(nsites=16) if name == "nacl" else S.argon(t_star=2.5))
    say("fused vv", name, s.n, cfg)
    out = []
    for fuse in (0, 1):
        e = E.Engine.from_system(s)
        e.set_launch(vv_fuse=fuse, refresh=0, **cfg)
        st = e.run(203, energy_every=7)
        st2 = e.run(50, energy_every=0)
        pos, vel, frc = e.download()
        out.append((st, st2, pos, vel, frc, e.info()["rebuilds"]))
        e.close()
    a, b = out
    say("  rebuilds", a[5], b[5])
    assert a[5] == b[5] and a[5] >= 5
    for k in range(5):
        if k < 2:
            for nm in a[k].dtype.names:
                assert np.array_equal(a[k][nm], b[k][nm]), nm
        else:
            assert np.array_equal(a[k], b[k])


def test_vv_thermostat_switch_starts_from_velocities(E):
    """Switching an NVE velocity-Verlet engine (leapfrog form) to Nose-Hoover
    restarts the thermostat from the current velocities: the run matches bit for
    bit an engine uploaded fresh with the same positions and velocities."""
    s = _nve_vv(S.argon())
    say("thermostat switch argon")
    e = E.Engine.from_system(s)
    e.set_launch(check_every=8)
    e.run(100, stats=False)
    pos, vel, _ = e.download()
    e.set_integrator("vverlet", "nosehoover", s.dt, 200.0)
    st = e.run(150)
    e.close()
    t = s.copy()
    t.pos, t.vel, t.thermostat, t.tau = pos, vel, "nosehoover", 200.0
    f = E.Engine.from_system(t)
    f.set_launch(check_every=8)
    st2 = f.run(150)
    f.close()
    say("  T", st["temp"][-1, 0], st2["temp"][-1, 0])
    for nm in ("ke", "temp", "conserved", "xi", "vxi"):
        assert np.array_equal(st[nm], st2[nm]), nm


def test_nve_conservation_nacl32768(E):
    """NaCl, 32,768 ions from the rocksalt lattice with 1200 K velocities, NVE
    leapfrog dt 2 fs, 1000 steps: the total energy (leapfrog half-step kinetic
    energy + potential) stays within 2 % of the kinetic energy.  (Half of the
    initial kinetic energy goes into the lattice's potential energy, so the
    temperature settles near 600 K.)"""
    s = S.nacl()
    s.thermostat = "none"
    say("nve nacl32768")
    e = E.Engine.from_system(s)
    e.set_launch(lanes=8, check_every=16)
    e.run(100, energy_every=0, stats=False)
    st = e.run(1000, energy_every=50)
    m = st["has_energy"][:, 0] == 1
    H = st["conserved"][m, 0]
    ke = st["ke"][m, 0]
    span = H.max() - H.min()
    say("  H span", span, "KE", ke.mean(), "T", st["temp"][-1, 0])
    assert np.all(np.isfinite(H))
    assert span < 0.02 * ke.mean()
    assert 300.0 < st["temp"][:, 0].min() and st["temp"][:, 0].max() < 1300.0
    e.close()


def test_invalid_calls_raise_and_leave_the_engine_usable(E):
    """Calls out of order, invalid launch fields, non-finite positions and a
    system beyond the slot limit return errors (MDError) instead of aborting,
    and the engine keeps working afterwards."""
    import ctypes as C
    s = S.argon(ncell=4)
    say("error paths")
    e = E.Engine()
    with pytest.raises(E.MDError):
        e.run(10)                                   # no force field, no system
    e.set_forcefield(s.ff)
    e.set_integrator(s.integrator, s.thermostat, s.dt, s.tau)
    n = (1 << 24) + 1                               # slot limit: 2^24 slots
    natom = np.array([n], np.int32)
    box = np.full(3, 1000.0, np.float32)
    T0 = np.array([100.0], np.float32)
    typ = np.zeros(n, np.int32)
    q = np.zeros(n, np.float32)
    m = np.full(n, 40.0, np.float32)
    rc = E.lib().md_set_system(e._h, 1, E._ptr(natom), E._ptr(box), E._ptr(T0), E._ptr(typ), E._ptr(q), E._ptr(m))
    assert rc == -4, rc                              # capacity
    del typ, q, m
    e.set_systems(s)
    for bad in (dict(lanes=3), dict(lanes=32), dict(cell_div=4), dict(check_every=0), dict(build_kernel=1),
                dict(defer=2)):
        with pytest.raises(E.MDError):
            e.set_launch(**bad)
    pos = s.pos.copy()
    pos[5, 1] = np.nan
    with pytest.raises(E.MDError):
        e.upload(pos)
    e.upload(s.pos, s.vel)
    st = e.run(50)
    assert np.all(np.isfinite(st["temp"])) and st["temp"][-1, 0] > 0
    e.close()


@pytest.mark.parametrize("kind", ["nacl_lf", "nacl_vv_nve", "water_vv_nh", "nacl_pme_lf", "water_pme_vv_nh"])
def test_list_overflow_rolls_back_and_replays(E, kind):
    """A list capacity below the rows' length (the reference's nmaxlist too
    small, which it does not guard) makes the run roll back to its start, grow
    the capacity and replay: the result equals a run that never overflowed.
    Leapfrog, fused NVE velocity Verlet and velocity Verlet + Nose-Hoover (the
    velocity-Verlet runs also restore the forces their first half kick reads),
    also with smooth PME (its grids are recomputed every step; the restored
    forces include the long-range part)."""
    if kind == "nacl_lf":
        s = S.nacl(nsites=16)
    elif kind == "nacl_vv_nve":
        s = dataclasses.replace(S.nacl(nsites=16), integrator="vverlet", thermostat="none")
    elif kind == "nacl_pme_lf":
        s = S.with_pme(S.nacl(nsites=16))
    elif kind == "water_pme_vv_nh":
        s = S.with_pme(S.water(nmol=512))
    else:
        s = S.water(nmol=512)                 # rows hold ~560 entries
    say("list overflow replay", kind, s.n)
    out = []
    for small in (False, True):
        e = E.Engine.from_system(s)
        e.set_launch(check_every=8)
        e.run(1, stats=False)                 # the run below starts at an odd step (the other position buffer)
        # both engines rebuild at the same positions (same atom order); NaCl rows hold ~410 entries
        e.set_list_capacity(128 if small else 1024)
        assert e.info()["cap"] == (128 if small else 1024)
        st = e.run(60, energy_every=20)
        pos, vel, frc = e.download()
        out.append((st, pos, vel, frc, e.info()["cap"]))
        e.close()
    say("  cap", out[0][4], out[1][4])
    assert out[1][4] > 128
    for nm in out[0][0].dtype.names:
        assert np.array_equal(out[0][0][nm], out[1][0][nm]), nm
    for k in (1, 2, 3):
        assert np.array_equal(out[0][k], out[1][k])
