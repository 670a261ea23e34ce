"""File formats of the drop-in boundary: the C++ SimulationInit parser
(src/simulation.cu:40-316 semantics) against the Python reader on the
reference's own example inputs, and write/read round trips."""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

from md_gpu_amd import fileio, systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_inputs")


class F4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class U4(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint), ("w", C.c_uint)]


class Config(C.Structure):
    _fields_ = [("natom", C.c_int), ("itype", C.POINTER(C.c_int)), ("box", F4), ("boxi", F4),
                ("pos", C.POINTER(F4)), ("vel", C.POINTER(F4)), ("force", C.POINTER(F4))]


class FField(C.Structure):
    _fields_ = [("ntype", C.c_int), ("nt", C.POINTER(C.c_int)), ("name", C.POINTER(C.c_char_p)),
                ("mass", C.POINTER(C.c_float)), ("charge", C.POINTER(C.c_float)), ("lj1", C.POINTER(C.c_float)),
                ("lj2", C.POINTER(C.c_float)), ("bnd1", C.POINTER(C.c_float)), ("bnd2", C.POINTER(C.c_float)),
                ("ang1", C.POINTER(C.c_float)), ("ang2", C.POINTER(C.c_float)), ("tbonds", C.c_int),
                ("tangle", C.c_int), ("nrigid", C.c_int), ("nbonds", C.c_int), ("nangle", C.c_int),
                ("nmaxbond", C.c_int), ("bndlist_max", C.c_void_p), ("bndlist", C.c_void_p),
                ("iatn", C.POINTER(U4)), ("anglist_max", C.c_void_p), ("anglist", C.c_void_p),
                ("iang", C.POINTER(U4))]


class Params(C.Structure):
    _fields_ = [("timeq", C.c_int), ("timrun", C.c_int), ("timprn", C.c_int), ("rcutsq", C.c_float),
                ("reacf", C.c_float), ("lkspace", C.c_int), ("kmax", C.c_int), ("alfa", C.c_float),
                ("skin", C.c_float), ("nmaxlist", C.c_int), ("rdel", C.c_float)]


class Integr(C.Structure):
    _fields_ = [("ensemble", C.c_char_p), ("enstype", C.c_int), ("dt", C.c_float), ("temper_0", C.c_float),
                ("tcons", C.c_float), ("press_0", C.c_float), ("pcons", C.c_float), ("temper", C.c_float),
                ("Ke_sum", C.c_void_p), ("Wi_sum", C.c_void_p), ("Ke_total", C.c_void_p), ("Wi_total", C.c_void_p),
                ("therm", C.c_void_p), ("integrator", C.c_int), ("thermostat", C.c_int)]


class Sim(C.Structure):
    _fields_ = [("simname", C.c_char_p), ("pars", C.POINTER(Params)), ("integ", C.POINTER(Integr)),
                ("conf", C.POINTER(Config)), ("fld", C.POINTER(FField))]


def compat():
    from md_gpu_amd import engine
    L = engine.lib()
    init = getattr(L, "_Z14SimulationInitPc")
    init.argtypes = [C.c_char_p]
    init.restype = C.POINTER(Sim)
    fin = getattr(L, "_Z16SimulationFinishP10simulation")
    fin.argtypes = [C.POINTER(Sim)]
    return init, fin


def stage(tmp_path, prefix, name):
    """Copy <prefix>.{inp,fld,cfg} fixtures to tmp as <name>.*"""
    for ext in ("inp", "fld", "cfg"):
        shutil.copy(os.path.join(GOLD, f"{prefix}.{ext}"), tmp_path / f"{name}.{ext}")
    return str(tmp_path / name)


def test_cpp_parser_matches_python_on_reference_h2o(tmp_path):
    name = stage(tmp_path, "h2o_wolf_h2o", "h2o")
    init, fin = compat()
    buf = C.create_string_buffer(name.encode(), 512)
    p = init(buf)
    assert p, "SimulationInit returned NULL"
    sim = p.contents
    pars, integ, conf, fld = sim.pars.contents, sim.integ.contents, sim.conf.contents, sim.fld.contents
    py = fileio.read_system(name)
    # .inp (src/simulation.cu:59-105): nvt, T0 300, 100000 steps, dt 1.0, Wolf rc 9 alpha 0.4, skin 0.5
    assert integ.enstype == 1 and integ.temper_0 == 300.0 and integ.dt == 1.0
    assert pars.timrun == 100000 and pars.timprn == 1000 and pars.lkspace == 1
    assert pars.rcutsq == np.float32(81.0) and pars.alfa == np.float32(0.4) and pars.skin == np.float32(0.5)
    assert py.ff.rc == 9.0 and py.ff.alpha == 0.4 and py.T0 == 300.0
    # .fld: masses x 1e4 (src/simulation.cu:177), LJ tables, bonds x2, angles deg->rad
    assert fld.ntype == 2
    assert [fld.name[i].decode() for i in range(2)] == py.ff.names
    assert np.allclose([fld.mass[i] for i in range(2)], 1e4 * py.ff.masses, rtol=1e-6)
    for i in range(2):
        for j in range(2):
            assert fld.lj1[i * 2 + j] == pytest.approx(py.ff.lj1[i, j], rel=1e-5)
            assert fld.lj2[i * 2 + j] == pytest.approx(py.ff.lj2[i, j], rel=1e-5)
    assert fld.tbonds == 1 and fld.bnd1[0] == 2 * py.ff.bond_k[0] and fld.bnd2[0] == np.float32(py.ff.bond_r0[0])
    assert fld.tangle == 1 and fld.ang1[0] == 2 * py.ff.angle_k[0]
    assert fld.ang2[0] == pytest.approx(py.ff.angle_t0[0], rel=1e-6)
    assert fld.nbonds == len(py.bonds) == 512 and fld.nangle == len(py.angles) == 256
    for k in (0, 1, 511):
        assert (fld.iatn[k].x - 1, fld.iatn[k].y - 1, fld.iatn[k].w - 1) == tuple(py.bonds[k])
    for k in (0, 255):
        a = fld.iang[k]
        assert (a.x - 1, a.y - 1, a.z - 1, a.w - 1) == tuple(py.angles[k])
    # .cfg: 768 atoms, charges per atom in pos.w, velocities (cftyp 1)
    assert conf.natom == py.n == 768
    assert conf.box.x == py.box[0]
    n = conf.natom
    pos = np.array([[conf.pos[i].x, conf.pos[i].y, conf.pos[i].z] for i in range(n)], np.float32)
    q = np.array([conf.pos[i].w for i in range(n)], np.float32)
    vel = np.array([[conf.vel[i].x, conf.vel[i].y, conf.vel[i].z] for i in range(n)], np.float32)
    typ = np.array([conf.itype[i] for i in range(n)])
    assert np.array_equal(pos, py.pos) and np.array_equal(q, py.charge) and np.array_equal(vel, py.vel)
    assert np.array_equal(typ, py.type)
    assert conf.vel[0].w == pytest.approx(1e4 * py.ff.masses[py.type[0]], rel=1e-6)
    # Finish writes <name>.xyz and <name>.xyz.cfg (src/simulation.cu:323,331)
    fin(p)
    assert os.path.exists(name + ".xyz") and os.path.exists(name + ".xyz.cfg")
    box, t2, q2, p2, v2 = fileio.read_cfg(name + ".xyz.cfg")
    assert np.allclose(p2, py.pos, atol=1e-6) and np.allclose(v2, py.vel, atol=1e-6) and np.array_equal(t2, py.type)
    lines = open(name + ".xyz").read().split("\n")
    assert lines[0] == "768" and lines[2].split()[0] == py.ff.names[py.type[0]]


def test_reference_inp_variants_parse(tmp_path):
    """.inp files of the reference examples (nve without a temperature line,
    reaction field with eps_rf, Wolf) parse with the C++ reader."""
    init, fin = compat()
    for prefix, ens, lk in (("wat_wat", 0, 0), ("nacl_nacl", 1, 0), ("ar_ar", 1, 0), ("co2_co2", None, None)):
        shutil.copy(os.path.join(GOLD, f"{prefix}.inp"), tmp_path / "x.inp")
        # pair it with a small generated force field / configuration
        s = S.argon(ncell=3)
        fileio.write_cfg(str(tmp_path / "x.cfg"), s.box, s.type, s.charge, s.pos, s.vel)
        with open(tmp_path / "x.fld", "w") as f:
            f.write("Atomtypes: 1\n1 108\nAr 39.948 0.0\nVDW_interactions: 1\n1 1 LJ 3.405 0.996\n"
                    "Bond_types: 0\nAngle_types: 0\nBonds: 0\nAngles: 0\n")
        buf = C.create_string_buffer(str(tmp_path / "x").encode(), 512)
        p = init(buf)
        assert p, prefix
        sim = p.contents
        if ens is not None:
            assert sim.integ.contents.enstype == ens and sim.pars.contents.lkspace == lk
        if prefix == "nacl_nacl":   # reacf: 1000.0 with rc 15 (src/simulation.cu:86)
            eps, rc = 1000.0, 15.0
            assert sim.pars.contents.reacf == pytest.approx(2 * (eps - 1) / (2 * eps + 1) / rc ** 3, rel=1e-6)
        fin(p)


def test_malformed_inputs_are_rejected(tmp_path):
    init, _ = compat()
    name = stage(tmp_path, "h2o_wolf_h2o", "bad")
    with open(name + ".cfg", "w") as f:
        f.write("768 1\n19.7 19.7 19.7\n1 2 0.41\n1.0 2.0\n")   # truncated
    assert not init(C.create_string_buffer(name.encode(), 512))
    assert not init(C.create_string_buffer(str(tmp_path / "missing").encode(), 512))


@pytest.mark.parametrize("mk", [lambda: S.water(nmol=64), lambda: S.nacl(nsites=6), lambda: S.argon(ncell=3)])
def test_write_read_round_trip(tmp_path, mk):
    s = mk()
    name = str(tmp_path / "sys")
    fileio.write_inputs(s, name)
    t = fileio.read_system(name)
    assert t.n == s.n and np.array_equal(t.type, s.type)
    assert np.allclose(t.pos, s.pos, atol=1e-6) and np.allclose(t.vel, s.vel, rtol=1e-6, atol=1e-9)
    assert np.allclose(t.charge, s.charge, atol=1e-6)
    assert np.allclose(t.ff.lj1, s.ff.lj1, rtol=1e-6) and np.allclose(t.ff.lj2, s.ff.lj2, rtol=1e-6)
    assert t.integrator == s.integrator and t.thermostat == s.thermostat
    assert np.array_equal(t.bonds, s.bonds) and np.array_equal(t.angles, s.angles)


def test_pme_key_round_trip_and_parser(tmp_path):
    """The optional "pme: K0 K1 K2 order" key (lkspace 2 only): written, read back,
    accepted by SimulationInit, refused after another lkspace."""
    s = S.with_pme(S.nacl(nsites=6), spacing=1.0, order=6)
    name = str(tmp_path / "pme")
    fileio.write_inputs(s, name)
    t = fileio.read_system(name)
    assert t.pme == s.pme and t.ff.coulomb == S.COUL_DAMPED
    assert t.ff.alpha == pytest.approx(3.09 / s.ff.rc)
    init, fin = compat()
    p = init(C.create_string_buffer(name.encode(), 512))
    assert p
    assert p.contents.pars.contents.lkspace == 2
    w = S.nacl(nsites=6)
    w.pme = s.pme   # Wolf (lkspace 1) with the key
    bad = str(tmp_path / "bad")
    fileio.write_inputs(w, bad)
    assert not init(C.create_string_buffer(bad.encode(), 512))
    fin(p)
