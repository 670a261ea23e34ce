"""ctypes wrapper of libmdgpu.so (the C ABI in include/mdgpu/md_api.h).

The wrapper mirrors the reference's module calls for the step path
(forces_init / config_init / neighborlist / forces / integrate, see
src/runGPU.cu:43-114) as methods of one Engine.  There is no CPU fallback:
loading fails loudly when the HIP library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence, Union

import numpy as np

from .systems import INTEGRATORS, THERMOSTATS, ForceField, System

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmdgpu.so")

MD_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "call out of order", -4: "capacity exceeded",
          -5: "unstable trajectory"}


class LaunchConfig(C.Structure):
    """md_launch_config (md_api.h)."""
    _fields_ = [("lanes", C.c_int), ("grid", C.c_int), ("cell_div", C.c_int), ("check_every", C.c_int),
                ("build_lanes", C.c_int), ("submit", C.c_int), ("refresh", C.c_int),
                ("build_kernel", C.c_int), ("defer", C.c_int), ("vv_fuse", C.c_int),
                ("fm_fused", C.c_int), ("persist", C.c_int)]

    @classmethod
    def names(cls):
        return [f[0] for f in cls._fields_]


class MDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{ERRORS.get(code, code)}] {msg}")
        self.code = code


class _FF(C.Structure):
    _fields_ = [("ntype", C.c_int), ("lj1", C.c_float * 64), ("lj2", C.c_float * 64), ("coulomb", C.c_int),
                ("rc", C.c_float), ("alpha", C.c_float), ("eps_rf", C.c_float), ("skin", C.c_float),
                ("qfac", C.c_float), ("nbond_types", C.c_int), ("bond_k2", C.c_float * 8),
                ("bond_r0", C.c_float * 8), ("nangle_types", C.c_int), ("angle_k2", C.c_float * 8),
                ("angle_t0", C.c_float * 8)]


class _Integ(C.Structure):
    _fields_ = [("integrator", C.c_int), ("thermostat", C.c_int), ("dt", C.c_float), ("tau", C.c_float),
                ("barostat", C.c_int), ("press0", C.c_float), ("tau_p", C.c_float), ("compress", C.c_float)]


BAROSTATS = {"none": 0, "berendsen": 1}


STAT_DTYPE = np.dtype([("ke", "f8"), ("temp", "f8"), ("pe_lj", "f8"), ("pe_coul", "f8"), ("pe_bond", "f8"),
                       ("pe_angle", "f8"), ("virial", "f8"), ("conserved", "f8"), ("xi", "f8"), ("vxi", "f8"),
                       ("volume", "f8"), ("has_energy", "i4"), ("pad", "i4")])

_lib = None


def lib():
    """Load libmdgpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not found: build it with `python -m md_gpu_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "md_last_error": ([], C.c_char_p), "md_version": ([], C.c_char_p),
            "md_create": ([C.POINTER(P), C.c_int], C.c_int), "md_destroy": ([P], C.c_int),
            "md_set_forcefield": ([P, P], C.c_int), "md_set_integrator": ([P, P], C.c_int),
            "md_set_system": ([P, C.c_int, P, P, P, P, P, P], C.c_int),
            "md_set_topology": ([P, C.c_int, P, C.c_int, P], C.c_int),
            "md_set_rigid": ([P, C.c_int, P, C.c_float, C.c_float], C.c_int),
            "md_set_pme": ([P, P, C.c_int], C.c_int),
            "md_upload_state": ([P, P, P], C.c_int), "md_download_state": ([P, P, P, P], C.c_int),
            "md_compute_forces": ([P, P, P, P], C.c_int), "md_run": ([P, C.c_int, C.c_int, P], C.c_int),
            "md_get_neighbors": ([P, C.c_int, P, P], C.c_int), "md_get_thresholds": ([P, P, P, P], C.c_int),
            "md_get_box": ([P, P], C.c_int), "md_get_sums": ([P, P, P], C.c_int),
            "md_check_lists": ([P, P], C.c_int),
            "md_get_info": ([P, P, P, P, P, P], C.c_int), "md_set_list_capacity": ([P, C.c_int], C.c_int),
            "md_set_launch": ([P, P], C.c_int), "md_get_launch": ([P, P], C.c_int),
            "md_autotune": ([P, C.c_int, P], C.c_int), "md_timing_enable": ([P, C.c_int], C.c_int),
            "md_timing_get": ([P, P, P, P, P], C.c_int), "md_copy_bench": ([P, C.c_longlong, C.c_int, P], C.c_int),
            "md_rebuild_bench": ([P, C.c_int, P], C.c_int), "md_persist_info": ([P, P, P, P], C.c_int),
            "md_fm_set_frames": ([P, C.c_int, P, P], C.c_int),
            "md_fm_loss_grad": ([P, P, P, C.c_float, P, P, P, P], C.c_int),
            "md_fm_loss_grad_noalloc": ([P, P, P, C.c_float, P, P, P, P], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def exported_symbols() -> List[str]:
    return ["md_last_error", "md_version", "md_create", "md_destroy", "md_set_forcefield", "md_set_integrator",
            "md_set_system", "md_set_topology", "md_upload_state", "md_download_state", "md_compute_forces",
            "md_run", "md_get_neighbors", "md_check_lists", "md_get_thresholds", "md_get_info", "md_set_launch",
            "md_get_launch", "md_autotune",
            "md_timing_enable", "md_timing_get", "md_copy_bench", "md_rebuild_bench", "md_fm_set_frames", "md_fm_loss_grad",
            "md_set_list_capacity", "md_fm_loss_grad_noalloc", "md_get_box", "md_get_sums", "md_persist_info",
            "md_set_rigid", "md_set_pme"]


def _check(rc: int):
    if rc != MD_OK:
        raise MDError(rc, lib().md_last_error().decode())


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def forcefield_struct(ff: ForceField, lj1=None, lj2=None) -> _FF:
    s = _FF()
    nt = ff.ntype
    s.ntype = nt
    a1 = np.zeros(64, np.float32)
    a2 = np.zeros(64, np.float32)
    a1[:nt * nt] = np.asarray(ff.lj1 if lj1 is None else lj1, np.float64).reshape(-1)
    a2[:nt * nt] = np.asarray(ff.lj2 if lj2 is None else lj2, np.float64).reshape(-1)
    s.lj1[:] = a1.tolist()
    s.lj2[:] = a2.tolist()
    s.coulomb = int(ff.coulomb)
    s.rc, s.alpha, s.eps_rf, s.skin, s.qfac = float(ff.rc), float(ff.alpha), float(ff.eps_rf), float(ff.skin), float(ff.qfac)
    bk = np.zeros(8, np.float32); br = np.zeros(8, np.float32); ak = np.zeros(8, np.float32); at = np.zeros(8, np.float32)
    nb, na = len(ff.bond_k), len(ff.angle_k)
    bk[:nb] = 2.0 * np.asarray(ff.bond_k); br[:nb] = ff.bond_r0
    ak[:na] = 2.0 * np.asarray(ff.angle_k); at[:na] = ff.angle_t0
    s.nbond_types, s.nangle_types = nb, na
    s.bond_k2[:] = bk.tolist(); s.bond_r0[:] = br.tolist(); s.angle_k2[:] = ak.tolist(); s.angle_t0[:] = at.tolist()
    return s


class Engine:
    """One MD engine on one GPU: R >= 1 replicas sharing a force field."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().md_create(C.byref(self._h), device))
        self.nrep = 0
        self.natom: List[int] = []
        self.N = 0

    def close(self):
        if self._h:
            lib().md_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up --------------------------------------------------------
    def set_forcefield(self, ff: ForceField):
        self._ff = ff
        s = forcefield_struct(ff)
        _check(lib().md_set_forcefield(self._h, C.byref(s)))

    def set_integrator(self, integrator: str, thermostat: str, dt: float, tau: float = 100.0,
                       barostat: str = "none", P0: float = 1.0, tau_p: float = 1000.0, compress: float = 4.5e-5):
        s = _Integ(INTEGRATORS[integrator], THERMOSTATS[thermostat], float(dt), float(tau), BAROSTATS[barostat],
                   float(P0), float(tau_p), float(compress))
        _check(lib().md_set_integrator(self._h, C.byref(s)))

    def set_systems(self, systems: Union[System, Sequence[System]]):
        """Set one or several replicas (same force field) and upload their state."""
        if isinstance(systems, System):
            systems = [systems]
        s0 = systems[0]
        self.set_forcefield(s0.ff)
        self.set_integrator(s0.integrator, s0.thermostat, s0.dt, s0.tau, s0.barostat, s0.P0, s0.tau_p, s0.compress)
        natom = np.array([s.n for s in systems], np.int32)
        box = np.concatenate([np.asarray(s.box, np.float32) for s in systems])
        T0 = np.array([s.T0 for s in systems], np.float32)
        typ = np.concatenate([s.type for s in systems]).astype(np.int32)
        q = np.concatenate([s.charge for s in systems]).astype(np.float32)
        m = np.concatenate([s.mass for s in systems]).astype(np.float32)
        _check(lib().md_set_system(self._h, len(systems), _ptr(natom), _ptr(box), _ptr(T0), _ptr(typ), _ptr(q), _ptr(m)))
        self.nrep = len(systems)
        self.natom = natom.tolist()
        self.N = int(natom.sum())
        offs = np.concatenate([[0], np.cumsum(natom)[:-1]])
        bonds = [s.bonds + np.array([o, o, 0], np.int32) for s, o in zip(systems, offs) if len(s.bonds)]
        angles = [s.angles + np.array([o, o, o, 0], np.int32) for s, o in zip(systems, offs) if len(s.angles)]
        b = np.concatenate(bonds).astype(np.int32) if bonds else np.zeros((0, 3), np.int32)
        a = np.concatenate(angles).astype(np.int32) if angles else np.zeros((0, 4), np.int32)
        _check(lib().md_set_topology(self._h, b.shape[0], _ptr(np.ascontiguousarray(b)), a.shape[0],
                                     _ptr(np.ascontiguousarray(a))))
        rig = [np.asarray(s.rigid, np.int32) + o for s, o in zip(systems, offs) if len(s.rigid)]
        if rig:
            self.set_rigid(np.concatenate(rig), *s0.rigid_d)
        pme = getattr(s0, "pme", ())
        if pme:
            if len(systems) != 1:
                raise MDError(-1, "PME runs one replica per engine")
            self.set_pme(pme[:3], pme[3])
        pos = np.ascontiguousarray(np.concatenate([s.pos for s in systems]), np.float32)
        vel = np.ascontiguousarray(np.concatenate([s.vel for s in systems]), np.float32)
        self.upload(pos, vel)

    def set_rigid(self, mol, d_oh: float, d_hh: float):
        """Rigid 3-site molecules (md_set_rigid): mol (nmol, 3) global atom indices (O, H1, H2)."""
        mol = np.ascontiguousarray(mol, np.int32).reshape(-1, 3)
        _check(lib().md_set_rigid(self._h, mol.shape[0], _ptr(mol), float(d_oh), float(d_hh)))

    def set_pme(self, grid, order: int):
        """Smooth PME for the damped Coulomb term (md_set_pme): grid (3,) points per
        box edge, B-spline order 4, 6 or 8; order 0 turns it off."""
        g = np.ascontiguousarray(grid, np.int32).reshape(3)
        _check(lib().md_set_pme(self._h, _ptr(g), int(order)))

    @classmethod
    def from_system(cls, systems, device: int = 0) -> "Engine":
        e = cls(device)
        e.set_systems(systems)
        return e

    # -- state ---------------------------------------------------------
    def upload(self, pos: np.ndarray, vel: Optional[np.ndarray] = None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        _check(lib().md_upload_state(self._h, _ptr(pos), _ptr(vel)))

    def download(self):
        pos = np.zeros((self.N, 3), np.float32)
        vel = np.zeros((self.N, 3), np.float32)
        frc = np.zeros((self.N, 3), np.float32)
        _check(lib().md_download_state(self._h, _ptr(pos), _ptr(vel), _ptr(frc)))
        return pos, vel, frc

    def compute_forces(self):
        F = np.zeros((self.N, 3), np.float32)
        E = np.zeros((self.nrep, 4), np.float64)
        W = np.zeros(self.nrep, np.float64)
        _check(lib().md_compute_forces(self._h, _ptr(F), _ptr(E), _ptr(W)))
        return F, E, W

    def run(self, nsteps: int, energy_every: int = 0, stats: bool = True):
        rec = np.zeros((nsteps, max(self.nrep, 1)), STAT_DTYPE) if stats else None
        _check(lib().md_run(self._h, int(nsteps), int(energy_every), _ptr(rec)))
        return rec

    def neighbors(self, cap: Optional[int] = None):
        if cap is None:
            cap = self.info()["cap"]
        counts = np.zeros(self.N, np.int32)
        nbrs = np.full((self.N, cap), -1, np.int32)
        _check(lib().md_get_neighbors(self._h, int(cap), _ptr(counts), _ptr(nbrs)))
        return counts, nbrs

    def check_lists(self):
        """Invariant violation counts of the device lists (md_check_lists): a dict
        of six counters, all zero when the sorted order and the lists are sound."""
        c = np.zeros(6, np.int32)
        _check(lib().md_check_lists(self._h, _ptr(c)))
        keys = ["order", "row_index", "row_distance", "set_entry", "capacity", "padding"]
        return dict(zip(keys, c.tolist()))

    def sums(self):
        """(force sum (3,), momentum sum (3,)) over all atoms, FP64, reduced on the device."""
        f = np.zeros(3)
        p = np.zeros(3)
        _check(lib().md_get_sums(self._h, _ptr(f), _ptr(p)))
        return f, p

    def box(self) -> np.ndarray:
        """(nrep, 3) current boxes (changed by pressure coupling)."""
        b = np.zeros(3 * max(self.nrep, 1), np.float32)
        _check(lib().md_get_box(self._h, _ptr(b)))
        return b.reshape(-1, 3)

    def thresholds(self):
        a = np.zeros(3, np.float32)
        _check(lib().md_get_thresholds(self._h, _ptr(a[0:1]), _ptr(a[1:2]), _ptr(a[2:3])))
        return float(a[0]), float(a[1]), float(a[2])

    def info(self):
        reb = C.c_longlong(); st = C.c_longlong(); cap = C.c_int(); ln = C.c_int(); gr = C.c_int()
        _check(lib().md_get_info(self._h, C.byref(reb), C.byref(st), C.byref(cap), C.byref(ln), C.byref(gr)))
        return dict(rebuilds=reb.value, steps=st.value, cap=cap.value, lanes=ln.value, grid=gr.value)

    # -- launch control / measurement ------------------------------------
    def set_list_capacity(self, cap: int):
        """Neighbor-list entries per atom (grown by the engine when a row overflows)."""
        _check(lib().md_set_list_capacity(self._h, int(cap)))

    def set_launch(self, **kw):
        """Change fields of the launch configuration (md_launch_config)."""
        c = LaunchConfig()
        _check(lib().md_get_launch(self._h, C.byref(c)))
        for k, v in kw.items():
            if k not in LaunchConfig.names():
                raise TypeError(f"unknown launch field {k}")
            setattr(c, k, int(v))
        _check(lib().md_set_launch(self._h, C.byref(c)))

    def get_launch(self):
        c = LaunchConfig()
        _check(lib().md_get_launch(self._h, C.byref(c)))
        return {k: getattr(c, k) for k in LaunchConfig.names()}

    def autotune(self, nsteps: int = 192) -> float:
        best = C.c_double()
        _check(lib().md_autotune(self._h, nsteps, C.byref(best)))
        return best.value

    def persist_info(self):
        """Runs made on the persistent small-system kernel, replays after a lane row
        overflow, and the current row length per lane (md_persist_info)."""
        runs = C.c_longlong(); rep = C.c_longlong(); rl = C.c_int()
        _check(lib().md_persist_info(self._h, C.byref(runs), C.byref(rep), C.byref(rl)))
        return dict(runs=runs.value, replays=rep.value, row_len=rl.value)

    def timing(self, enable: bool):
        _check(lib().md_timing_enable(self._h, 1 if enable else 0))

    def timing_get(self):
        a = C.c_double(); b = C.c_longlong(); c = C.c_double(); d = C.c_longlong()
        _check(lib().md_timing_get(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(step_ms=a.value, step_launches=b.value, rebuild_ms=c.value, rebuilds=d.value)

    def rebuild_bench(self, reps: int = 10) -> float:
        """Mean ms of one forced full list rebuild (all replicas)."""
        ms = C.c_double()
        _check(lib().md_rebuild_bench(self._h, reps, C.byref(ms)))
        return ms.value

    def copy_bench(self, nbytes: int, reps: int = 20) -> float:
        ms = C.c_double()
        _check(lib().md_copy_bench(self._h, C.c_longlong(nbytes), reps, C.byref(ms)))
        return ms.value

    # -- force matching ------------------------------------------------
    def fm_set_frames(self, frames: np.ndarray, targets: np.ndarray):
        frames = np.ascontiguousarray(frames, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        nf = frames.shape[0]
        _check(lib().md_fm_set_frames(self._h, nf, _ptr(frames), _ptr(targets)))
        self.nrep = nf
        self.N = frames.shape[0] * frames.shape[1]

    def fm_loss_grad(self, lj1: np.ndarray, lj2: np.ndarray, qscale: float, noalloc: bool = False):
        """(loss, g_lj1, g_lj2, g_qscale); noalloc=True goes through
        md_fm_loss_grad_noalloc (the library's fixed device scratch)."""
        nt = lj1.shape[0]
        a1 = np.ascontiguousarray(lj1, np.float32)
        a2 = np.ascontiguousarray(lj2, np.float32)
        loss = C.c_double(); gq = C.c_double()
        g1 = np.zeros((nt, nt)); g2 = np.zeros((nt, nt))
        fn = lib().md_fm_loss_grad_noalloc if noalloc else lib().md_fm_loss_grad
        _check(fn(self._h, _ptr(a1), _ptr(a2), C.c_float(qscale), C.byref(loss), _ptr(g1), _ptr(g2), C.byref(gq)))
        return loss.value, g1, g2, gq.value


def version() -> str:
    return lib().md_version().decode()
