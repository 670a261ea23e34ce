"""Reader and writer of the reference's input files (.inp, .fld, .cfg) and of
its .xyz / .xyz.cfg output, for System records.

Formats (src/simulation.cu:40-345): whitespace-separated, positional, labels
ignored; 1-based atom and type indices; masses in g/mol; LJ pairs as
"i j LJ sigma epsilon" (or "12-6 A B"); bond / angle types as "idx k x0" with
E = k (x - x0)^2 and angles in degrees; .cfg per atom "idx type charge" then
"x y z" (and "vx vy vz" when cftyp > 0).
"""
from __future__ import annotations

import math
import os
from typing import Tuple

import numpy as np

from .systems import COUL_DAMPED, COUL_NONE, COUL_RF, COUL_WOLF, ForceField, System, eps_sig_from_lj


def write_inputs(s: System, name: str, timrun: int = 100, timprn: int = 10, eps_rf: float = 78.0):
    """Write <name>.inp, <name>.fld, <name>.cfg for System s."""
    ff = s.ff
    ens = "npt" if s.barostat != "none" else ("nve" if s.thermostat == "none" else "nvt")
    lk = {COUL_RF: 0, COUL_WOLF: 1, COUL_DAMPED: 2, COUL_NONE: 1}[ff.coulomb]
    with open(name + ".inp", "w") as f:
        f.write(f"ensemble: {ens}\n")
        if ens == "nvt":
            f.write(f"temper: {s.T0:.6f}\n")
        elif ens == "npt":   # temperature and pressure (bar), src/simulation.cu:71-75
            f.write(f"temper_press: {s.T0:.6f} {s.P0:.6f}\n")
        f.write(f"times: {timrun} {timrun} {timprn} {s.dt}\n")
        f.write(f"cutoff: {ff.rc} {lk}\n")
        if lk == 0:
            f.write(f"reacf: {ff.eps_rf if ff.coulomb == COUL_RF else eps_rf}\n")
        elif lk == 1:
            f.write(f"alfa: {ff.alpha}\n")
        else:
            f.write("kmax: 0\n")
        f.write(f"skin: {ff.skin} 1000\n")
        f.write("rdel: 0.02\n")
        f.write(f"integrator: {s.integrator}\n")
        f.write(f"thermostat: {s.thermostat} {s.tau}\n")
        if ens == "npt":
            f.write(f"barostat: {s.barostat} {s.tau_p} {s.compress}\n")
        if getattr(s, "pme", ()):
            f.write("pme: {} {} {} {}\n".format(*s.pme))
    nt = ff.ntype
    counts = np.bincount(s.type, minlength=nt)
    qtype = [float(s.charge[s.type == t][0]) if counts[t] else 0.0 for t in range(nt)]
    eps, sig = eps_sig_from_lj(ff.lj1, ff.lj2)
    pairs = [(i, j) for i in range(nt) for j in range(i, nt) if ff.lj1[i, j] != 0 or ff.lj2[i, j] != 0]
    with open(name + ".fld", "w") as f:
        f.write(f"Atomtypes: {nt}\n")
        for t in range(nt):
            f.write(f"{t + 1} {counts[t]}\n{ff.names[t]} {ff.masses[t]:.6f} {qtype[t]:.6f}\n")
        f.write(f"\nVDW_interactions: {len(pairs)}\n")
        for i, j in pairs:
            f.write(f"{i + 1} {j + 1}  LJ  {sig[i, j]:.9g}  {eps[i, j]:.9g}\n")
        f.write(f"\nBond_types: {len(ff.bond_k)}\n")
        for k, (kb, r0) in enumerate(zip(ff.bond_k, ff.bond_r0)):
            f.write(f"{k + 1} {kb:.9g} {r0:.9g}\n")
        f.write(f"\nAngle_types: {len(ff.angle_k)}\n")
        for k, (ka, t0) in enumerate(zip(ff.angle_k, ff.angle_t0)):
            f.write(f"{k + 1} {ka:.9g} {math.degrees(t0):.9g}\n")
        f.write(f"\nBonds: {len(s.bonds)}\n")
        for k, (i, j, t) in enumerate(s.bonds):
            f.write(f"{k + 1} {t + 1} {i + 1} {j + 1}\n")
        f.write(f"\nAngles: {len(s.angles)}\n")
        for k, (i, j, m, t) in enumerate(s.angles):
            f.write(f"{k + 1} {t + 1} {i + 1} {j + 1} {m + 1}\n")
    write_cfg(name + ".cfg", s.box, s.type, s.charge, s.pos, s.vel)


def write_cfg(path, box, type_, charge, pos, vel=None):
    lines = [f"{len(pos)} {1 if vel is not None else 0}", f"{box[0]:.6f} {box[1]:.6f} {box[2]:.6f}"]
    for i in range(len(pos)):
        lines.append(f"{i + 1} {int(type_[i]) + 1} {float(charge[i]):.6f}")
        lines.append(f"{pos[i, 0]:.6f} {pos[i, 1]:.6f} {pos[i, 2]:.6f}")
        if vel is not None:
            lines.append(f"{vel[i, 0]:.9g} {vel[i, 1]:.9g} {vel[i, 2]:.9g}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_cfg(path) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """(box, type 0-based, charge, pos, vel or None) of a .cfg / .xyz.cfg file."""
    tok = open(path).read().split()
    n, cftyp = int(tok[0]), int(tok[1])
    box = np.array([float(x) for x in tok[2:5]], np.float32)
    per = 6 + (3 if cftyp > 0 else 0)
    a = tok[5:5 + n * per]
    typ = np.array([int(a[k * per + 1]) - 1 for k in range(n)], np.int32)
    q = np.array([float(a[k * per + 2]) for k in range(n)], np.float32)
    pos = np.array([[float(a[k * per + 3 + d]) for d in range(3)] for k in range(n)], np.float32)
    vel = None
    if cftyp > 0:
        vel = np.array([[float(a[k * per + 6 + d]) for d in range(3)] for k in range(n)], np.float32)
    return box, typ, q, pos, vel


def read_fld(path):
    """Parse a .fld into a dict (reference order, src/simulation.cu:144-264)."""
    tok = open(path).read().split()
    p = 0

    def nxt():
        nonlocal p
        p += 1
        return tok[p - 1]
    nxt()
    nt = int(nxt())
    names, masses, charges, counts = [], [], [], []
    for _ in range(nt):
        nxt(); counts.append(int(nxt()))
        names.append(nxt()); masses.append(float(nxt())); charges.append(float(nxt()))
    nxt()
    nv = int(nxt())
    lj1 = np.zeros((nt, nt)); lj2 = np.zeros((nt, nt))
    for _ in range(nv):
        i, j, form, a, b = int(nxt()) - 1, int(nxt()) - 1, nxt(), float(nxt()), float(nxt())
        if form == "LJ":
            v1, v2 = 48.0 * b * a ** 12, 24.0 * b * a ** 6
        else:
            v1, v2 = 12.0 * a, 6.0 * b
        lj1[i, j] = lj1[j, i] = v1
        lj2[i, j] = lj2[j, i] = v2
    nxt(); nbt = int(nxt())
    bk, br = [], []
    for _ in range(nbt):
        nxt(); bk.append(float(nxt())); br.append(float(nxt()))
    nxt(); nat = int(nxt())
    ak, at = [], []
    for _ in range(nat):
        nxt(); ak.append(float(nxt())); at.append(math.radians(float(nxt())))
    nxt(); nb = int(nxt())
    bonds = []
    for _ in range(nb):
        nxt(); t = int(nxt()) - 1; i = int(nxt()) - 1; j = int(nxt()) - 1
        bonds.append((i, j, t))
    nxt(); na = int(nxt())
    angles = []
    for _ in range(na):
        nxt(); t = int(nxt()) - 1; i = int(nxt()) - 1; j = int(nxt()) - 1; k = int(nxt()) - 1
        angles.append((i, j, k, t))
    return dict(names=names, masses=np.array(masses), charges=np.array(charges), counts=counts, lj1=lj1, lj2=lj2,
                bond_k=np.array(bk), bond_r0=np.array(br), angle_k=np.array(ak), angle_t0=np.array(at),
                bonds=np.array(bonds, np.int32).reshape(-1, 3), angles=np.array(angles, np.int32).reshape(-1, 4))


def read_system(name: str) -> System:
    """System from <name>.inp/.fld/.cfg (the inverse of write_inputs)."""
    tok = open(name + ".inp").read().split()
    p = 1
    ens = tok[p]; p += 1
    T0, P0 = 300.0, 1.0
    if ens == "nvt":
        T0 = float(tok[p + 1]); p += 2
    elif ens == "npt":
        T0, P0 = float(tok[p + 1]), float(tok[p + 2]); p += 3
    dt = float(tok[p + 4]); p += 5
    rc = float(tok[p + 1]); lk = int(tok[p + 2]); p += 3
    alpha, eps_rf = 0.0, 1.0
    if lk == 0:
        eps_rf = float(tok[p + 1])
    elif lk == 1:
        alpha = float(tok[p + 1])
    else:
        alpha = 3.09 / rc
    p += 2
    skin = float(tok[p + 1]); p += 3
    p += 2
    integ, thermo, tau = "leapfrog", ("berendsen" if ens in ("nvt", "npt") else "none"), 100.0
    baro, tau_p, compress = ("berendsen" if ens == "npt" else "none"), 1000.0, 4.5e-5
    pme = ()
    while p < len(tok):
        if tok[p] == "integrator:":
            integ = tok[p + 1]; p += 2
        elif tok[p] == "thermostat:":
            thermo, tau = tok[p + 1], float(tok[p + 2]); p += 3
        elif tok[p] == "barostat:":
            baro, tau_p, compress = tok[p + 1], float(tok[p + 2]), float(tok[p + 3]); p += 4
        elif tok[p] == "pme:":
            pme = tuple(int(x) for x in tok[p + 1:p + 5]); p += 5
        else:
            raise ValueError(f"unknown key {tok[p]}")
    fld = read_fld(name + ".fld")
    box, typ, q, pos, vel = read_cfg(name + ".cfg")
    coul = {0: COUL_RF, 1: COUL_WOLF, 2: COUL_DAMPED}[lk]
    ff = ForceField(fld["names"], fld["masses"], fld["lj1"], fld["lj2"], coulomb=coul, rc=rc, alpha=alpha,
                    eps_rf=eps_rf, skin=skin, bond_k=fld["bond_k"], bond_r0=fld["bond_r0"],
                    angle_k=fld["angle_k"], angle_t0=fld["angle_t0"])
    if vel is None:
        vel = np.zeros_like(pos)
    return System(os.path.basename(name), box, pos, vel, typ, q, ff, bonds=fld["bonds"], angles=fld["angles"],
                  integrator=integ, thermostat=thermo if ens != "nve" else "none", dt=dt, T0=T0, tau=tau,
                  barostat=baro if ens == "npt" else "none", P0=P0, tau_p=tau_p, compress=compress, pme=pme)
