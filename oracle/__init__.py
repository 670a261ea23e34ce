"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
md_gpu_amd never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

MAXT = 8


class OrcFF(C.Structure):
    _fields_ = [("ntype", C.c_int), ("lj1", C.c_double * 64), ("lj2", C.c_double * 64),
                ("coulomb", C.c_int), ("rc", C.c_double), ("alpha", C.c_double), ("eps_rf", C.c_double),
                ("qfac", C.c_double), ("erfc_exact", C.c_int),
                ("nbtype", C.c_int), ("bk2", C.c_double * 8), ("br0", C.c_double * 8),
                ("natype", C.c_int), ("ak2", C.c_double * 8), ("at0", C.c_double * 8)]


class OrcMD(C.Structure):
    _fields_ = [("integrator", C.c_int), ("thermostat", C.c_int), ("dt", C.c_double), ("T0", C.c_double),
                ("tau", C.c_double), ("skin", C.c_double), ("barostat", C.c_int), ("P0", C.c_double),
                ("tau_p", C.c_double), ("compress", C.c_double), ("nrigid", C.c_int), ("rigid", C.c_void_p),
                ("d_oh", C.c_double), ("d_hh", C.c_double), ("pme_order", C.c_int), ("pme_K", C.c_int * 3)]


class OrcPME(C.Structure):
    _fields_ = [("K", C.c_int * 3), ("order", C.c_int), ("beta", C.c_double)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        L.orc_canonical_thresholds.argtypes = [C.c_float, C.c_float, P, P, P, P, P]
        L.orc_rsq32.argtypes = [P, P, P, P]
        L.orc_rsq32.restype = C.c_float
        for f in (L.orc_nlist_brute, L.orc_nlist_cells):
            f.argtypes = [C.c_int, P, P, C.c_float, P, C.c_int, P, P]
            f.restype = C.c_int
        L.orc_pair_forces.argtypes = [C.c_int, P, P, P, P, P, C.c_float, C.c_int, P, P, P, P, P]
        L.orc_bond_forces.argtypes = [C.c_int, P, P, P, C.c_int, P, P, P, P]
        L.orc_angle_forces.argtypes = [C.c_int, P, P, P, C.c_int, P, P, P]
        L.orc_md_run.argtypes = [C.c_int, P, P, P, P, P, P, P, P, C.c_int, P, C.c_int, P, P, C.c_int,
                                 P, P, P, P, P, P, P]
        L.orc_md_run.restype = C.c_int
        L.orc_fm_loss_grad.argtypes = [C.c_int, C.c_int, P, P, P, P, P, C.c_double, P, P, P, P]
        L.orc_settle.argtypes = [C.c_int, P, P, P, C.c_double, C.c_double, P, P]
        L.orc_settle.restype = None
        L.orc_shake.argtypes = [C.c_int, P, P, P, C.c_double, C.c_double, P, P, C.c_double, C.c_int]
        L.orc_shake.restype = C.c_int
        L.orc_fm_loss_grad.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        L.orc_fft3.argtypes = [P, P, C.c_int]
        L.orc_fft3.restype = None
        L.orc_bspline.argtypes = [C.c_int, C.c_double, P, P]
        L.orc_bspline.restype = None
        L.orc_pme_recip.argtypes = [P, C.c_int, P, P, P, C.c_double, P, P]
        L.orc_ewald_recip.argtypes = [C.c_int, P, P, P, C.c_double, C.c_double, C.c_int, P, P]
        L.orc_pme_excl.argtypes = [C.c_int, P, P, P, P, C.c_double, C.c_double, P, P]
        L.orc_ewald_self.argtypes = [C.c_int, P, P, C.c_double, C.c_double, P]
        for f in (L.orc_pme_recip, L.orc_ewald_recip, L.orc_pme_excl, L.orc_ewald_self):
            f.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_ff(ff, erfc_exact: int = 0, lj1=None, lj2=None) -> OrcFF:
    """OrcFF from an md_gpu_amd.systems.ForceField (double precision copy)."""
    o = OrcFF()
    nt = ff.ntype
    o.ntype = nt
    l1 = np.zeros(64)
    l2 = np.zeros(64)
    l1[:nt * nt] = np.asarray(ff.lj1 if lj1 is None else lj1, np.float64).reshape(-1)
    l2[:nt * nt] = np.asarray(ff.lj2 if lj2 is None else lj2, np.float64).reshape(-1)
    o.lj1[:] = l1.tolist()
    o.lj2[:] = l2.tolist()
    o.coulomb = int(ff.coulomb)
    o.rc = float(ff.rc)
    o.alpha = float(ff.alpha)
    o.eps_rf = float(ff.eps_rf)
    o.qfac = float(ff.qfac)
    o.erfc_exact = int(erfc_exact)
    bk = np.zeros(8); br = np.zeros(8); ak = np.zeros(8); at = np.zeros(8)
    nb = len(ff.bond_k); na = len(ff.angle_k)
    bk[:nb] = 2.0 * np.asarray(ff.bond_k); br[:nb] = ff.bond_r0
    ak[:na] = 2.0 * np.asarray(ff.angle_k); at[:na] = ff.angle_t0
    o.nbtype, o.natype = nb, na
    o.bk2[:] = bk.tolist(); o.br0[:] = br.tolist(); o.ak2[:] = ak.tolist(); o.at0[:] = at.tolist()
    return o


def thresholds(rc: float, skin: float, box):
    box = np.ascontiguousarray(box, np.float32)
    rc2 = np.zeros(1, np.float32); rl2 = np.zeros(1, np.float32); s2 = np.zeros(1, np.float32)
    linv = np.zeros(3, np.float32)
    lib().orc_canonical_thresholds(C.c_float(rc), C.c_float(skin), _p(box), _p(rc2), _p(rl2), _p(s2), _p(linv))
    return float(rc2[0]), float(rl2[0]), float(s2[0]), linv


def rsq32(xi, xj, box):
    box = np.ascontiguousarray(box, np.float32)
    linv = (1.0 / box.astype(np.float64)).astype(np.float32)
    xi = np.ascontiguousarray(xi, np.float32); xj = np.ascontiguousarray(xj, np.float32)
    return lib().orc_rsq32(_p(xi), _p(xj), _p(box), _p(linv))


def nlist(pos, box, rl2, excl=None, cap=None, method="cells"):
    """Full neighbor lists as a list of sorted int arrays."""
    pos = np.ascontiguousarray(pos, np.float32)
    box = np.ascontiguousarray(box, np.float32)
    n = pos.shape[0]
    ex = None if excl is None else np.ascontiguousarray(excl, np.int32)
    f = lib().orc_nlist_cells if method == "cells" else lib().orc_nlist_brute
    counts = np.zeros(n, np.int32)
    if cap is None:
        dummy = np.zeros(n, np.int32)
        cap = max(1, f(n, _p(pos), _p(box), C.c_float(rl2), _p(ex), 1, _p(counts), _p(dummy)))
    nb = np.zeros((n, cap), np.int32)
    f(n, _p(pos), _p(box), C.c_float(rl2), _p(ex), cap, _p(counts), _p(nb))
    return counts, nb


def pair_forces(sys_, counts, nbrs, rc2=None, erfc_exact=0, want_energy=True, charge=None, ff=None):
    """FP64 pair forces (n,3), energies [lj, coul], virial."""
    ff = ff or sys_.ff
    pos = np.ascontiguousarray(sys_.pos, np.float32)
    box = np.ascontiguousarray(sys_.box, np.float32)
    typ = np.ascontiguousarray(sys_.type, np.int32)
    q = np.ascontiguousarray(sys_.charge if charge is None else charge, np.float64)
    if rc2 is None:
        rc2 = thresholds(ff.rc, ff.skin, box)[0]
    o = make_ff(ff, erfc_exact)
    F = np.zeros((sys_.n, 3))
    e = np.zeros(2)
    w = np.zeros(1)
    counts = np.ascontiguousarray(counts, np.int32)
    nbrs = np.ascontiguousarray(nbrs, np.int32)
    lib().orc_pair_forces(sys_.n, _p(pos), _p(typ), _p(q), _p(box), C.byref(o), C.c_float(rc2), nbrs.shape[1],
                          _p(counts), _p(nbrs), _p(F), _p(e) if want_energy else None,
                          _p(w) if want_energy else None)
    return F, e, float(w[0])


def bonded_forces(sys_, pos=None):
    """FP64 bond + angle forces (n,3) and energies [bond, angle], bond virial."""
    p = np.ascontiguousarray(sys_.pos if pos is None else pos, np.float32)
    box = np.ascontiguousarray(sys_.box, np.float32)
    o = make_ff(sys_.ff)
    F = np.zeros((sys_.n, 3))
    eb = np.zeros(1); ea = np.zeros(1); wb = np.zeros(1)
    b = np.ascontiguousarray(sys_.bonds, np.int32)
    a = np.ascontiguousarray(sys_.angles, np.int32)
    lib().orc_bond_forces(sys_.n, _p(p), _p(box), C.byref(o), b.shape[0], _p(b), _p(F), _p(eb), _p(wb))
    lib().orc_angle_forces(sys_.n, _p(p), _p(box), C.byref(o), a.shape[0], _p(a), _p(F), _p(ea))
    return F, np.array([eb[0], ea[0]]), float(wb[0])


def forces(sys_, erfc_exact=0, cap=None):
    """Total FP64 force with canonical FP32 list/cutoff predicates.
    Returns F (n,3), energies [lj, coul, bond, angle], virial (pair + bond).
    With sys_.pme the Coulomb energy, forces and virial include the Ewald
    reciprocal part, the excluded-pair correction and the self term."""
    rc2, rl2, _, _ = thresholds(sys_.ff.rc, sys_.ff.skin, sys_.box)
    counts, nb = nlist(sys_.pos, sys_.box, rl2, sys_.exclusions(), cap)
    Fp, ep, wp = pair_forces(sys_, counts, nb, rc2, erfc_exact)
    Fb, eb, wb = bonded_forces(sys_)
    if getattr(sys_, "pme", ()):
        F, e, w = ewald_long_range(sys_, sys_.pos)
        Fp = Fp + F
        ep = ep.copy()
        ep[1] += e
        wp = wp + w
    return Fp + Fb, np.concatenate([ep, eb]), wp + wb


def _pme_par(K, order, beta):
    o = OrcPME()
    o.K[:] = [int(k) for k in K]
    o.order = int(order)
    o.beta = float(beta)
    return o


def pme_recip(x, q, box, K, order, beta, qfac):
    """SPME reciprocal (energy, forces (n, 3), virial trace) in FP64."""
    x = np.ascontiguousarray(x, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    b = np.ascontiguousarray(box, np.float32)
    F = np.zeros_like(x)
    W = np.zeros(1)
    E = lib().orc_pme_recip(C.byref(_pme_par(K, order, beta)), x.shape[0], _p(x), _p(q), _p(b), float(qfac),
                            _p(F), _p(W))
    return E, F, float(W[0])


def ewald_recip(x, q, box, beta, qfac, kmax):
    """Direct Ewald reciprocal sum over |m_d| <= kmax: (energy, forces, virial trace)."""
    x = np.ascontiguousarray(x, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    b = np.ascontiguousarray(box, np.float32)
    F = np.zeros_like(x)
    W = np.zeros(1)
    E = lib().orc_ewald_recip(x.shape[0], _p(x), _p(q), _p(b), float(beta), float(qfac), int(kmax), _p(F), _p(W))
    return E, F, float(W[0])


def pme_excl(x, q, box, excl, beta, qfac):
    """Excluded-pair correction -qq erf(beta r)/r: (energy, forces, virial)."""
    x = np.ascontiguousarray(x, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    b = np.ascontiguousarray(box, np.float32)
    ex = np.ascontiguousarray(excl, np.int32)
    F = np.zeros_like(x)
    W = np.zeros(1)
    E = lib().orc_pme_excl(x.shape[0], _p(x), _p(q), _p(b), _p(ex), float(beta), float(qfac), _p(F), _p(W))
    return E, F, float(W[0])


def ewald_self(q, box, beta, qfac):
    """Self plus neutralizing-background term: (energy, virial)."""
    q = np.ascontiguousarray(q, np.float64)
    b = np.ascontiguousarray(box, np.float32)
    W = np.zeros(1)
    E = lib().orc_ewald_self(q.shape[0], _p(q), _p(b), float(beta), float(qfac), _p(W))
    return E, float(W[0])


def ewald_long_range(sys_, x):
    """Everything the damped real-space sum leaves to the rest of Ewald for a system
    with sys_.pme = (K0, K1, K2, order): reciprocal + excluded pairs + self."""
    K0, K1, K2, order = sys_.pme
    beta, qf = sys_.ff.alpha, sys_.ff.qfac
    q = np.asarray(sys_.charge, np.float64)
    e1, F1, w1 = pme_recip(x, q, sys_.box, (K0, K1, K2), order, beta, qf)
    e2, F2, w2 = pme_excl(x, q, sys_.box, sys_.exclusions(), beta, qf)
    e3, w3 = ewald_self(q, sys_.box, beta, qf)
    return F1 + F2, e1 + e2 + e3, w1 + w2 + w3


def fft3(a, sign=-1):
    """Unnormalized 3D DFT of a complex array (the oracle's own FFT)."""
    z = np.ascontiguousarray(a, np.complex128).copy()
    K = np.ascontiguousarray(z.shape, np.int32)
    lib().orc_fft3(_p(z), _p(K), int(sign))
    return z


def bspline(order, w):
    """(M_p(w + j), M_p'(w + j)) for j = 0..p-1."""
    th = np.zeros(order)
    dth = np.zeros(order)
    lib().orc_bspline(int(order), float(w), _p(th), _p(dth))
    return th, dth


def md_run(sys_, nsteps, pos=None, vel=None, thermo_state=None, skin=None):
    """FP64 trajectory.  Returns dict of per-step arrays and the final state."""
    n = sys_.n
    p = np.ascontiguousarray(sys_.pos if pos is None else pos, np.float64).copy()
    v = np.ascontiguousarray(sys_.vel if vel is None else vel, np.float64).copy()
    typ = np.ascontiguousarray(sys_.type, np.int32)
    q = np.ascontiguousarray(sys_.charge, np.float64)
    m = np.ascontiguousarray(sys_.mass, np.float64)
    box = np.ascontiguousarray(sys_.box, np.float32)
    o = make_ff(sys_.ff)
    md = OrcMD()
    md.integrator = {"leapfrog": 0, "vverlet": 1}[sys_.integrator]
    md.thermostat = {"none": 0, "berendsen": 1, "nosehoover": 2}[sys_.thermostat]
    md.dt, md.T0, md.tau = float(sys_.dt), float(sys_.T0), float(sys_.tau)
    md.skin = float(sys_.ff.skin if skin is None else skin)
    md.barostat = {"none": 0, "berendsen": 1}[getattr(sys_, "barostat", "none")]
    md.P0 = float(getattr(sys_, "P0", 1.0))
    md.tau_p = float(getattr(sys_, "tau_p", 1000.0))
    md.compress = float(getattr(sys_, "compress", 4.5e-5))
    rig = np.ascontiguousarray(getattr(sys_, "rigid", np.zeros((0, 3), np.int32)), np.int32)
    md.nrigid = int(rig.shape[0])
    md.rigid = rig.ctypes.data if rig.shape[0] else None
    md.d_oh, md.d_hh = (float(x) for x in getattr(sys_, "rigid_d", (1.0, 1.0)))
    pme = getattr(sys_, "pme", ())
    if pme:
        md.pme_K[:] = [int(k) for k in pme[:3]]
        md.pme_order = int(pme[3])
    b = np.ascontiguousarray(sys_.bonds, np.int32)
    a = np.ascontiguousarray(sys_.angles, np.int32)
    ex = np.ascontiguousarray(sys_.exclusions(), np.int32)
    ke = np.zeros(nsteps); pe = np.zeros(nsteps); T = np.zeros(nsteps); H = np.zeros(nsteps)
    ts = np.zeros(2) if thermo_state is None else np.ascontiguousarray(thermo_state, np.float64).copy()
    vol = np.zeros(nsteps)
    box_end = np.zeros(3)
    nreb = lib().orc_md_run(n, _p(p), _p(v), _p(typ), _p(q), _p(m), _p(box), C.byref(o), C.byref(md),
                            b.shape[0], _p(b), a.shape[0], _p(a), _p(ex), nsteps, _p(ke), _p(pe), _p(T), _p(H),
                            _p(ts), _p(vol), _p(box_end))
    if nreb == -1:
        raise ValueError("the oracle runs rigid molecules with leapfrog only")
    if nreb == -2:
        raise ValueError("PME needs the damped Coulomb term")
    return dict(ke=ke, pe=pe, temp=T, conserved=H, pos=p, vel=v, thermo=ts, rebuilds=nreb, volume=vol, box=box_end)


def settle(mol, mass, box, d_oh, d_hh, x0, x1):
    """SETTLE: x1 (n, 3) constrained from the positions x0 before the step; returns a new array."""
    mol = np.ascontiguousarray(mol, np.int32)
    m = np.ascontiguousarray(mass, np.float64)
    b = np.ascontiguousarray(box, np.float32)
    a0 = np.ascontiguousarray(x0, np.float64)
    a1 = np.ascontiguousarray(x1, np.float64).copy()
    lib().orc_settle(mol.shape[0], _p(mol), _p(m), _p(b), float(d_oh), float(d_hh), _p(a0), _p(a1))
    return a1


def shake(mol, mass, box, d_oh, d_hh, x0, x1, tol=1e-14, maxit=10000):
    """SHAKE on the same constraints: (constrained copy of x1, most iterations used or -1)."""
    mol = np.ascontiguousarray(mol, np.int32)
    m = np.ascontiguousarray(mass, np.float64)
    b = np.ascontiguousarray(box, np.float32)
    a0 = np.ascontiguousarray(x0, np.float64)
    a1 = np.ascontiguousarray(x1, np.float64).copy()
    it = lib().orc_shake(mol.shape[0], _p(mol), _p(m), _p(b), float(d_oh), float(d_hh), _p(a0), _p(a1),
                         float(tol), int(maxit))
    return a1, it


def fm_loss_grad(frames, type_, charge, box, ff, qscale, targets, lj1=None, lj2=None):
    """FP64 force-matching loss and gradient (g_lj1 (nt,nt), g_lj2 (nt,nt), g_qscale)."""
    frames = np.ascontiguousarray(frames, np.float32)
    nf, n = frames.shape[0], frames.shape[1]
    o = make_ff(ff, 0, lj1, lj2)
    nt = ff.ntype
    g1 = np.zeros(nt * nt); g2 = np.zeros(nt * nt); gq = np.zeros(1)
    typ = np.ascontiguousarray(type_, np.int32)
    q = np.ascontiguousarray(charge, np.float64)
    b = np.ascontiguousarray(box, np.float32)
    t = np.ascontiguousarray(targets, np.float64)
    loss = lib().orc_fm_loss_grad(nf, n, _p(frames), _p(typ), _p(q), _p(b), C.byref(o), C.c_double(qscale), _p(t),
                                  _p(g1), _p(g2), _p(gq))
    return loss, g1.reshape(nt, nt), g2.reshape(nt, nt), float(gq[0])


def num_threads() -> int:
    return int(lib().orc_num_threads())
