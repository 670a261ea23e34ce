"""Seeded generators for the benchmark and test systems, plus the force-field record.

Units follow the reference throughout (src/simulation.cu:177, src/forces.cu:47):
angstrom, femtosecond, kJ/mol, elementary charge, g/mol.  Internal mass is
1e4 * m so that F/m_int is in A/fs^2.  Temperature uses the reference's
T = sum(m_int v^2) / ((3N-3) * 8.314e-3) (src/integrator.cu:153-155).

LJ parameters are stored the way the reference stores them after reading a
.fld "LJ" line (src/simulation.cu:191-195): lj1 = 48 eps sig^12 and
lj2 = 24 eps sig^6, so that F/r = lj1 r^-14 - lj2 r^-8 (src/pair_forces.cu:190-193).
Bond and angle constants are kept as k of E = k (x - x0)^2; the engine's
bnd1 = 2k / ang1 = 2k factors follow src/simulation.cu:227,242.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

QFAC = 1389.35427611722          # src/forces.cu:47, kJ A / (mol e^2)
KB_REF = 8.314e-3                # temperature conversion of src/integrator.cu:155
KB_KJ = 1.3806503e-23 * 6.0221415e23 / 1000.0   # K -> kJ/mol, src/simulation.h:35-36,189

COUL_RF, COUL_WOLF, COUL_DAMPED, COUL_NONE = 0, 1, 2, 3
INTEGRATORS = {"leapfrog": 0, "vverlet": 1}
THERMOSTATS = {"none": 0, "berendsen": 1, "nosehoover": 2}


def lj_pair(eps: float, sig: float):
    """(lj1, lj2) for one type pair, src/simulation.cu:191,194."""
    return 48.0 * eps * sig ** 12, 24.0 * eps * sig ** 6


@dataclasses.dataclass
class ForceField:
    names: List[str]
    masses: np.ndarray            # (ntype,) g/mol
    lj1: np.ndarray               # (ntype, ntype)
    lj2: np.ndarray
    coulomb: int = COUL_WOLF
    rc: float = 10.0
    alpha: float = 0.2
    eps_rf: float = 1.0
    skin: float = 1.0
    bond_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    bond_r0: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    angle_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    angle_t0: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))   # radians
    qfac: float = QFAC

    @property
    def ntype(self) -> int:
        return len(self.names)

    @staticmethod
    def from_eps_sig(names, masses, eps, sig, **kw) -> "ForceField":
        """eps, sig: (ntype, ntype) matrices (0 = no LJ)."""
        eps = np.asarray(eps, dtype=np.float64)
        sig = np.asarray(sig, dtype=np.float64)
        lj1 = 48.0 * eps * sig ** 12
        lj2 = 24.0 * eps * sig ** 6
        return ForceField(list(names), np.asarray(masses, np.float64), lj1, lj2, **kw)


@dataclasses.dataclass
class System:
    name: str
    box: np.ndarray               # (3,) float32
    pos: np.ndarray               # (n, 3) float32, box centred at the origin
    vel: np.ndarray               # (n, 3) float32 (leapfrog: v(t - dt/2))
    type: np.ndarray              # (n,) int32, 0-based
    charge: np.ndarray            # (n,) float32
    ff: ForceField
    bonds: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3), np.int32))
    angles: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 4), np.int32))
    integrator: str = "leapfrog"
    thermostat: str = "none"
    dt: float = 1.0
    T0: float = 300.0
    tau: float = 100.0            # fs; Berendsen tco = 0.01 /fs in src/integrator.cu:146
    barostat: str = "none"        # "berendsen": isotropic Berendsen pressure coupling (leapfrog)
    P0: float = 1.0               # bar
    tau_p: float = 1000.0         # fs
    compress: float = 4.5e-5      # isothermal compressibility, 1/bar
    # rigid 3-site molecules (leapfrog only): atom indices (O, H1, H2) per molecule, held
    # at the distances rigid_d = (d_OH, d_HH) by SETTLE after every step
    rigid: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3), np.int32))
    rigid_d: tuple = (1.0, 1.6330)
    # smooth particle-mesh Ewald for the reciprocal part of the damped (lkspace 2)
    # Coulomb term: (K0, K1, K2, order), empty = none
    pme: tuple = ()

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def mass(self) -> np.ndarray:
        return self.ff.masses[self.type].astype(np.float64)

    def exclusions(self) -> np.ndarray:
        """(n, 4) int32 1-2 bonded partners, -1 = empty (src/neighborlist.cu:64-99,
        filled correctly: the reference's slot test at :89-94 reads excl[i]); the
        atoms of a rigid molecule exclude each other."""
        ex = np.full((self.n, 4), -1, np.int32)
        fill = np.zeros(self.n, np.int32)
        pairs = [(int(i), int(j)) for i, j, _ in self.bonds]
        for o, h1, h2 in self.rigid:
            pairs += [(int(o), int(h1)), (int(o), int(h2)), (int(h1), int(h2))]
        for i, j in pairs:
            for a, b in ((i, j), (j, i)):
                if b in ex[a, :fill[a]]:
                    continue
                if fill[a] >= 4:
                    raise ValueError(f"atom {a} has more than 4 excluded partners")
                ex[a, fill[a]] = b
                fill[a] += 1
        return ex

    def nf(self) -> int:
        """Degrees of freedom of the temperature: 3N - 3 (src/integrator.cu:155), less
        three per rigid molecule."""
        return 3 * self.n - 3 - 3 * int(self.rigid.shape[0])

    def temperature(self, vel: Optional[np.ndarray] = None) -> float:
        v = self.vel if vel is None else vel
        k2 = float(np.sum(1e4 * self.mass[:, None] * v.astype(np.float64) ** 2))
        return k2 / (self.nf() * KB_REF)

    def copy(self) -> "System":
        return dataclasses.replace(self, box=self.box.copy(), pos=self.pos.copy(), vel=self.vel.copy(),
                                   type=self.type.copy(), charge=self.charge.copy(),
                                   bonds=self.bonds.copy(), angles=self.angles.copy(), rigid=self.rigid.copy())


# ---------------------------------------------------------------- helpers

def maxwell_boltzmann(rng: np.random.Generator, mass: np.ndarray, T: float) -> np.ndarray:
    """Gaussian velocities at T, centre-of-mass momentum removed, rescaled so
    that the reference temperature (N_f = 3N - 3) is exactly T."""
    m = 1e4 * np.asarray(mass, np.float64)
    v = rng.standard_normal((m.size, 3)) * np.sqrt(KB_REF * T / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    t_now = float((m[:, None] * v ** 2).sum()) / ((3 * m.size - 3) * KB_REF)
    if t_now > 0:
        v *= math.sqrt(T / t_now)
    return v


def wrap_centered(pos: np.ndarray, box: np.ndarray) -> np.ndarray:
    return pos - box * np.floor(pos / box + 0.5)


def fcc_sites(m: int, a: float) -> np.ndarray:
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    return ((g[:, None, :] + basis[None]) * a).reshape(-1, 3)


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


# ---------------------------------------------------------------- workloads

def argon(ncell: int = 6, seed: int = 1, rho_star: float = 0.8442, t_star: float = 0.722) -> System:
    """Plain LJ argon on an FCC lattice (ncell^3 unit cells, 4 atoms each), NVE velocity Verlet, dt 2 fs."""
    sig, eps_k, mass = 3.405, 119.8, 39.948
    eps = eps_k * KB_KJ
    a = (4.0 / (rho_star / sig ** 3)) ** (1.0 / 3.0)
    L = ncell * a
    pos = fcc_sites(ncell, a) + 0.25 * a - 0.5 * L
    n = pos.shape[0]
    ff = ForceField.from_eps_sig(["Ar"], [mass], [[eps]], [[sig]], coulomb=COUL_NONE,
                                 rc=2.5 * sig, skin=0.3 * sig)
    rng = np.random.default_rng(seed)
    typ = np.zeros(n, np.int32)
    vel = maxwell_boltzmann(rng, ff.masses[typ], t_star * eps_k)
    box = np.full(3, L, np.float32)
    return System("argon", box, wrap_centered(pos, L).astype(np.float32), vel.astype(np.float32), typ,
                  np.zeros(n, np.float32), ff, integrator="vverlet", thermostat="none", dt=2.0,
                  T0=t_star * eps_k)


def fft_size(n: int) -> int:
    """Smallest m >= n whose prime factors are 2, 3 and 5."""
    m = max(int(n), 1)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def with_pme(s: System, spacing: float = 1.0, order: int = 4, alpha: Optional[float] = None) -> System:
    """s with Ewald electrostatics: the damped real-space term (lkspace 2,
    src/forces.cu:97-110) with alpha = 3.09 / rc by default (the reference's
    estimate, src/simulation.cu:93) plus the smooth-PME reciprocal part on a grid
    of at most `spacing` A per point and B-splines of `order`."""
    t = s.copy()
    t.ff = dataclasses.replace(s.ff, coulomb=COUL_DAMPED, alpha=float(3.09 / s.ff.rc if alpha is None else alpha))
    K = [fft_size(int(math.ceil(float(L) / spacing))) for L in s.box]
    t.pme = (K[0], K[1], K[2], int(order))
    t.name = s.name + "_pme"
    return t


def nacl_ff(alpha=0.2, rc=12.0, skin=1.0) -> ForceField:
    """Na/Cl LJ with Lorentz-Berthelot mixing written into explicit pair entries
    (the reference has no mixing rule, src/simulation.cu:184-205), Wolf Coulomb."""
    sig = np.array([2.35, 4.40])
    eps = np.array([0.13, 0.42])
    S = 0.5 * (sig[:, None] + sig[None, :])
    E = np.sqrt(eps[:, None] * eps[None, :])
    return ForceField.from_eps_sig(["Na", "Cl"], [22.98977, 35.4527], E, S, coulomb=COUL_WOLF,
                                   rc=rc, alpha=alpha, skin=skin)


def nacl(nsites: int = 32, seed: int = 2, T: float = 1200.0, a: float = 5.64, ff: Optional[ForceField] = None) -> System:
    """Rocksalt NaCl: nsites^3 ions on a simple-cubic grid of spacing a/2 with
    alternating charge (nsites even).  Default = 32,768 ions, L = 90.24 A,
    NVT leapfrog + Berendsen tau 0.1 ps, dt 2 fs."""
    if nsites % 2:
        raise ValueError("nsites must be even")
    h = a / 2.0
    L = nsites * h
    g = np.stack(np.meshgrid(np.arange(nsites), np.arange(nsites), np.arange(nsites), indexing="ij"), -1).reshape(-1, 3)
    pos = g * h + 0.25 * h - 0.5 * L
    typ = (g.sum(1) % 2).astype(np.int32)             # 0 = Na+, 1 = Cl-
    q = np.where(typ == 0, 1.0, -1.0).astype(np.float32)
    ff = ff or nacl_ff()
    rng = np.random.default_rng(seed)
    vel = maxwell_boltzmann(rng, ff.masses[typ], T)
    box = np.full(3, L, np.float32)
    return System("nacl", box, wrap_centered(pos, L).astype(np.float32), vel.astype(np.float32), typ, q, ff,
                  integrator="leapfrog", thermostat="berendsen", dt=2.0, T0=T, tau=100.0)


def water(nmol: int = 65536, seed: int = 3, T: float = 300.0, density: float = 1.0, rigid: bool = False) -> System:
    """Flexible 3-site water at density g/cm^3, O on a cubic lattice (random
    subset of sites), random orientations; harmonic O-H and H-O-H, LJ on O only,
    charges O -0.82 / H +0.41, Wolf alpha 0.25 rc 10, NVT velocity Verlet
    Nose-Hoover 300 K, dt 0.5 fs.  Exclusions are 1-2 only (reference semantics).

    rigid=True: the same molecules held rigid by SETTLE (O-H 1 A, H-O-H 109.47
    degrees, so H-H 1.633 A), no bonds or angles, the three atoms of a molecule
    excluding each other; NVT leapfrog Berendsen (tau 100 fs), dt 2 fs."""
    m_o, m_h = 15.9994, 1.0079
    mw = m_o + 2 * m_h
    nd = density * 6.0221415e23 / mw * 1e-24          # molecules / A^3
    L = (nmol / nd) ** (1.0 / 3.0)
    k = int(math.ceil(nmol ** (1.0 / 3.0) - 1e-9))
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)
    pick = np.sort(rng.choice(g.shape[0], nmol, replace=False))
    sites = (g[pick] + 0.5) * (L / k) - 0.5 * L
    r0, th0 = 1.0, math.radians(109.47)
    local = np.array([[0, 0, 0],
                      [r0 * math.sin(th0 / 2), 0, r0 * math.cos(th0 / 2)],
                      [-r0 * math.sin(th0 / 2), 0, r0 * math.cos(th0 / 2)]])
    R = random_rotations(rng, nmol)
    pos = (sites[:, None, :] + np.einsum("mij,aj->mai", R, local)).reshape(-1, 3)
    typ = np.tile(np.array([0, 1, 1], np.int32), nmol)
    q = np.tile(np.array([-0.82, 0.41, 0.41], np.float32), nmol)
    ff = ForceField.from_eps_sig(["O", "H"], [m_o, m_h], [[0.650, 0], [0, 0]], [[3.166, 0], [0, 0]],
                                 coulomb=COUL_WOLF, rc=10.0, alpha=0.25, skin=1.0,
                                 bond_k=np.array([4637.0]), bond_r0=np.array([1.0]),
                                 angle_k=np.array([383.0]), angle_t0=np.array([th0]))
    base = 3 * np.arange(nmol, dtype=np.int32)
    bonds = np.concatenate([np.stack([base, base + 1, np.zeros_like(base)], 1),
                            np.stack([base, base + 2, np.zeros_like(base)], 1)]).astype(np.int32)
    angles = np.stack([base + 1, base, base + 2, np.zeros_like(base)], 1).astype(np.int32)
    vel = maxwell_boltzmann(rng, ff.masses[typ], T)
    box = np.full(3, L, np.float32)
    if rigid:
        mol = np.stack([base, base + 1, base + 2], 1).astype(np.int32)
        return System("water_rigid", box, wrap_centered(pos, L).astype(np.float32), vel.astype(np.float32), typ, q,
                      ff, integrator="leapfrog", thermostat="berendsen", dt=2.0, T0=T, tau=100.0, rigid=mol,
                      rigid_d=(r0, 2.0 * r0 * math.sin(th0 / 2)))
    return System("water", box, wrap_centered(pos, L).astype(np.float32), vel.astype(np.float32), typ, q, ff,
                  bonds=bonds, angles=angles, integrator="vverlet", thermostat="nosehoover", dt=0.5, T0=T,
                  tau=100.0)


def arxe_ff(skin: float = 0.3 * 4.10) -> ForceField:
    """Ar-Xe Lennard-Jones with Lorentz-Berthelot mixing, cutoff 2.5 sigma_max and
    skin 0.3 sigma_max (the skin rule of the argon configuration, 0.3 sigma)."""
    sig = np.array([3.405, 4.10])
    eps = np.array([119.8, 221.0]) * KB_KJ
    S = 0.5 * (sig[:, None] + sig[None, :])
    E = np.sqrt(eps[:, None] * eps[None, :])
    return ForceField.from_eps_sig(["Ar", "Xe"], [39.948, 131.293], E, S, coulomb=COUL_NONE,
                                   rc=2.5 * 4.10, skin=skin)


def arxe_replicas(nrep: int = 512, seed: int = 4, nmin: int = 256, nmax: int = 8192,
                  rho: float = 0.018, xe_frac: float = 0.3) -> List[System]:
    """Independent Ar-Xe LJ mixtures: sizes log-uniform in [nmin, nmax], per-replica
    T uniform in 80-300 K, number density rho (A^-3), 30% Xe; NVT leapfrog Berendsen, dt 2 fs."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), nrep))).astype(int)
    temps = rng.uniform(80.0, 300.0, nrep)
    ff = arxe_ff()
    out = []
    for r in range(nrep):
        n = int(sizes[r])
        L = (n / rho) ** (1.0 / 3.0)
        m = int(math.ceil((n / 4.0) ** (1.0 / 3.0) - 1e-9))
        sites = fcc_sites(m, L / m) + 0.25 * (L / m) - 0.5 * L
        pick = np.sort(rng.choice(sites.shape[0], n, replace=False))
        pos = sites[pick]
        typ = np.zeros(n, np.int32)
        typ[rng.choice(n, int(round(xe_frac * n)), replace=False)] = 1
        vel = maxwell_boltzmann(rng, ff.masses[typ], temps[r])
        box = np.full(3, L, np.float32)
        out.append(System(f"arxe{r}", box, wrap_centered(pos, L).astype(np.float32), vel.astype(np.float32), typ,
                          np.zeros(n, np.float32), ff, integrator="leapfrog", thermostat="berendsen", dt=2.0,
                          T0=float(temps[r]), tau=100.0))
    return out


def perturb_nacl_ff(ff: ForceField, seed: int, frac: float = 0.05):
    """Target model for force matching: eps, sigma of each type pair and the
    charge scale multiplied by 1 + U(-frac, frac).  Returns (ff_target, qscale_target)."""
    rng = np.random.default_rng(seed)
    nt = ff.ntype
    eps, sig = eps_sig_from_lj(ff.lj1, ff.lj2)
    iu = np.triu_indices(nt)
    fe = 1.0 + rng.uniform(-frac, frac, len(iu[0]))
    fs = 1.0 + rng.uniform(-frac, frac, len(iu[0]))
    E, S = eps.copy(), sig.copy()
    E[iu] *= fe
    S[iu] *= fs
    E[(iu[1], iu[0])] = E[iu]
    S[(iu[1], iu[0])] = S[iu]
    qs = 1.0 + rng.uniform(-frac, frac)
    tgt = dataclasses.replace(ff, lj1=48.0 * E * S ** 12, lj2=24.0 * E * S ** 6)
    return tgt, float(qs)


def eps_sig_from_lj(lj1: np.ndarray, lj2: np.ndarray):
    """Invert lj1 = 48 eps sig^12, lj2 = 24 eps sig^6, i.e. sig^6 = lj1 / (2 lj2),
    eps = lj2 / (24 sig^6) (entries with lj2 = 0 give 0)."""
    lj1 = np.asarray(lj1, np.float64)
    lj2 = np.asarray(lj2, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        s6 = np.where(lj2 > 0, lj1 / (2.0 * np.where(lj2 > 0, lj2, 1.0)), 0.0)
        sig = np.where(s6 > 0, s6 ** (1.0 / 6.0), 0.0)
        eps = np.where(s6 > 0, lj2 / (24.0 * np.where(s6 > 0, s6, 1.0)), 0.0)
    return eps, sig
