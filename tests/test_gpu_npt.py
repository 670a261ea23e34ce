"""NPT on the GPU: leapfrog with Berendsen temperature and pressure coupling
(isotropic) against the FP64 oracle's barostat, which the CPU tests pin to
the ideal-gas volume, the NVT limit and the pressure response.

The oracle rebuilds its lists every coupled step; the engine keeps its lists
through the box scaling (its skin check measures each atom's move from its
scaled build position against (min scale x r_list - r_c) / 2), so agreement
also shows that check keeps every pair within the cutoff listed."""
import dataclasses
import os
import subprocess
import sys

import numpy as np
import pytest

from md_gpu_amd import fileio, systems as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print("[case]", *a, file=sys.stderr, flush=True)


def npt(s, **kw):
    kw.setdefault("tau_p", 200.0)
    kw.setdefault("compress", 4.5e-4)
    return dataclasses.replace(s, integrator="leapfrog", thermostat="berendsen", tau=100.0, barostat="berendsen",
                               **kw)


@pytest.mark.parametrize("P0", [1.0, 5000.0])
def test_npt_argon_matches_oracle(orc, P0):
    from md_gpu_amd import engine as E
    s = npt(S.argon(), P0=P0)
    n = 60
    say("npt argon", s.n, "P0", P0)
    e = E.Engine.from_system(s)
    pos0, vel0, _ = e.download()
    so = s.copy()
    so.pos, so.vel = pos0, vel0
    st = e.run(n, energy_every=1)
    pos, vel, _ = e.download()
    box = e.box()[0].astype(np.float64)
    bad = e.check_lists()
    r = orc.md_run(so, n)
    say("  box gpu", box, "oracle", r["box"], "V0", np.prod(s.box), "rebuilds", e.info()["rebuilds"], bad)
    assert all(v == 0 for v in bad.values()), bad
    assert abs(np.prod(r["box"]) / np.prod(s.box) - 1.0) > 1e-3          # the box did move
    np.testing.assert_allclose(box, r["box"], rtol=2e-5)
    np.testing.assert_allclose(st["volume"][:, 0], r["volume"], rtol=2e-5)
    np.testing.assert_allclose(st["temp"][:, 0], r["temp"], rtol=2e-4)
    d = pos.astype(np.float64) - r["pos"]
    d -= box * np.rint(d / box)
    assert np.abs(d).max() < 5e-4
    e.close()


def test_npt_batched_replicas_match_oracle(orc):
    """Three Ar-Xe replicas, each with its own coupled box, in one engine."""
    from md_gpu_amd import engine as E
    reps = [npt(x, P0=2000.0) for x in S.arxe_replicas(nrep=3, seed=12, nmin=600, nmax=1500)]
    n = 40
    say("npt replicas", [x.n for x in reps])
    e = E.Engine.from_system(reps)
    st = e.run(n, energy_every=1)
    pos, _, _ = e.download()
    boxes = e.box().astype(np.float64)
    off = 0
    for k, x in enumerate(reps):
        r = orc.md_run(x, n)
        say("  replica", k, "box gpu", boxes[k], "oracle", r["box"])
        np.testing.assert_allclose(boxes[k], r["box"], rtol=2e-5)
        np.testing.assert_allclose(st["volume"][:, k], r["volume"], rtol=2e-5)
        np.testing.assert_allclose(st["temp"][:, k], r["temp"], rtol=2e-4)
        d = pos[off:off + x.n].astype(np.float64) - r["pos"]
        d -= boxes[k] * np.rint(d / boxes[k])
        assert np.abs(d).max() < 5e-4
        off += x.n
    e.close()


def test_athom_npt_writes_the_coupled_box(tmp_path, orc):
    """The reference's NPT input (ensemble npt, temperature and pressure) through
    athom: the restart file holds the coupled box and positions."""
    s = npt(S.nacl(nsites=12), P0=3000.0)
    name = str(tmp_path / "nacl_npt")
    fileio.write_inputs(s, name, timrun=20, timprn=10)
    t = fileio.read_system(name)
    assert t.barostat == "berendsen" and t.P0 == pytest.approx(3000.0)
    say("athom npt", s.n)
    r = subprocess.run([os.path.join(ROOT, "md_gpu_amd", "bin", "athom"), name], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    ref = orc.md_run(t, 21)
    box, typ, q, pos, vel = fileio.read_cfg(name + ".xyz.cfg")
    say("  box", box, "oracle", ref["box"])
    np.testing.assert_allclose(box, ref["box"], rtol=2e-5)
    d = pos.astype(np.float64) - ref["pos"]
    d -= box * np.rint(d / box)
    assert np.abs(d).max() < 5e-4
