"""Known-answer test of the oracle against forces the reference logged.

tests/golden/nacl_rf_step0 holds the reference's examples/nacl/nacl.cfg (4096
ions, box 100 A) and the step-0 forces its run printed (examples/nacl/x) for
the ions i % 128 == 0, the rows the logged build reduced correctly (first
thread of each 128-thread block).  The run used reaction field with
eps_rf = 1000, rc = 15 (examples/nacl/nacl.inp) and one LJ pair for all type
pairs with lj1, lj2 from the log header.  Pins the LJ + reaction-field force
law of src/pair_forces.cu:34-131 (the r^-14 / r^-8 split, the 1/r^3 - reacf
form, qfac, the cutoff) on data the reference produced."""
import os

import numpy as np
import pytest

from md_gpu_amd import fileio, systems as S

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nacl_rf_step0")


def load_fixture():
    box, typ, q, pos, vel = fileio.read_cfg(os.path.join(HERE, "nacl.cfg"))
    with open(os.path.join(HERE, "forces.txt")) as f:
        hdr = [l.split() for l in f if l.startswith("#")]
    consts = dict(zip(hdr[1][1::2], map(float, hdr[1][2::2])))
    rows = np.loadtxt(os.path.join(HERE, "forces.txt"), comments="#")
    idx = rows[:, 0].astype(int)
    ff = S.ForceField(names=["Na", "Cl"], masses=np.array([22.98977, 35.4527]),
                      lj1=np.full((2, 2), consts["lj1"]), lj2=np.full((2, 2), consts["lj2"]),
                      coulomb=S.COUL_RF, rc=15.0, alpha=0.0, eps_rf=1000.0, skin=0.4)
    s = S.System(name="nacl_rf_step0", box=box.astype(np.float32), pos=pos.astype(np.float32),
                 vel=vel.astype(np.float32), type=typ, charge=q.astype(np.float32), ff=ff)
    return s, consts, idx, rows[:, 1:4]


def test_fixture_constants_are_the_input_files():
    s, consts, idx, _ = load_fixture()
    assert s.n == 4096 and np.array_equal(idx, np.arange(0, 4096, 128))
    assert consts["rc2"] == 15.0 ** 2
    # logged reacf = 2 (eps - 1) / ((2 eps + 1) rc^3) at eps_rf = 1000, rc = 15 (src/simulation.cu:86)
    assert consts["reacf"] == pytest.approx(2 * 999 / (2001 * 15.0 ** 3), abs=5e-7)


def test_oracle_reproduces_logged_forces(orc):
    s, _, idx, F_ref = load_fixture()
    F, _, _ = orc.forces(s)
    rel = np.linalg.norm(F[idx] - F_ref, axis=1) / np.linalg.norm(F_ref, axis=1)
    assert rel.max() < 3e-5, rel.max()


def test_fixture_discriminates_force_law(orc):
    """The rows separate the reaction-field law from plain cut Coulomb (reacf = 0)."""
    s, _, idx, F_ref = load_fixture()
    s.ff.eps_rf = 1.0
    F, _, _ = orc.forces(s)
    rel = np.linalg.norm(F[idx] - F_ref, axis=1) / np.linalg.norm(F_ref, axis=1)
    assert np.median(rel) > 1e-3
