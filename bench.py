#!/usr/bin/env python3
"""Benchmark of the MI355X MD step engine.

One JSON line on stdout.  The headline `value` is timesteps/s of the largest
workload: 512 independent Ar-Xe LJ replicas (log-uniform 256..8192 atoms,
1.29 M atoms in all, per-replica T 80..300 K, rc 10.25 A, NVT leapfrog +
Berendsen, dt 2 fs, FP32 forces) advanced together, one step-kernel launch
per step, inputs resident in HBM, hipGraph segments.  The other workloads
(water 196,608 atoms NVT-VV Nose-Hoover, molten NaCl 32,768 ions NVT-LF
Berendsen, argon 864 NVE-VV, NaCl force matching) are reported under
"workloads" with the same fields.

Roofline of the dominant kernel (the fused step kernel): algorithmic FP32
flops per launch = f_pair * P (P = unordered pairs within rc at the timed
configuration, f_pair counted from src/pair_forces.cu:151-200: LJ+Wolf 65,
Wolf only 60, LJ 36) divided by its mean launch time measured with HIP events
on the engine's stream, against 157.3 TFLOP/s.  `traffic` is the HBM bytes per
launch of the same kernel from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary
under profiles/.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3            # TFLOP/s, MI355X
PEAK_HBM = 8000.0            # GB/s
METRIC = "timesteps/s (ns/day) + force-grad evals/s at ≤1e-5 force err, 1 MI355X"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def count_pairs(pos, box, rc, types=None, tmask=None):
    """Unordered pairs with r < rc (periodic), optionally split by a type predicate."""
    from scipy.spatial import cKDTree
    p = np.mod(pos.astype(np.float64) + box / 2, box)
    p = np.minimum(p, np.nextafter(box, 0))
    tree = cKDTree(p, boxsize=box.astype(np.float64))
    pairs = tree.query_pairs(rc, output_type="ndarray")
    if types is None:
        return len(pairs), None
    t = types[pairs[:, 0]], types[pairs[:, 1]]
    return len(pairs), int(np.sum(tmask(*t)))


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def time_run(e, nsteps, energy_every=0):
    t0 = time.perf_counter()
    e.run(nsteps, energy_every=energy_every, stats=False)
    return time.perf_counter() - t0


def traffic_of(workload, kernel="step_kernel"):
    """HBM bytes per launch of the workload's dominant kernel from the counter summary in profiles/ (or None).
    The persistent kernel runs a whole run in one launch: its bytes are given per step."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        tr = json.load(f)
    return tr.get("bytes_per_step", tr.get("bytes_per_launch")), os.path.relpath(path, ROOT)


def kernel_time(e, nsteps):
    """Mean duration of the fused step kernel and of a rebuild, HIP events on the engine stream."""
    e.timing(True)
    e.run(nsteps, energy_every=0, stats=False)
    t = e.timing_get()
    e.timing(False)
    return t["step_ms"] / max(t["step_launches"], 1), t["rebuild_ms"] / max(t["rebuilds"], 1), t


def step_workload(name, systems, f_pair_fn, steps, warmup, autotune, dt, cpu_fn=None, force_check=None,
                  launch=None, calibrate=False, long_steps=0):
    from md_gpu_amd import engine as E
    log(f"{name}: setting up")
    e = E.Engine.from_system(systems)
    tuned = None
    if launch:
        e.set_launch(**launch)
        log(f"{name}: launch {e.get_launch()}")
    elif autotune:
        tuned = e.autotune(steps)   # tuned for runs of the timed run's length
        log(f"{name}: autotuned {e.get_launch()} ({tuned:.4f} ms/step)")
    pos0, _, _ = e.download()
    flops0 = f_pair_fn(pos0)
    # the warm-up runs right before the timed run: host work in between (pair
    # counting takes up to seconds) would let the GPU clocks drop
    e.run(warmup, stats=False)
    reb0 = e.info()["rebuilds"]
    wall = time_run(e, steps)
    reb1 = e.info()["rebuilds"]
    pos1, _, _ = e.download()
    flops1 = f_pair_fn(pos1)
    flops = 0.5 * (flops0 + flops1)
    long_ms = None
    if long_steps:   # the same engine over a long run: per-step time in steady state
        long_ms = 1e3 * time_run(e, long_steps) / long_steps
    ktime_steps = max(20, min(steps, 2000))
    kstep_ms, kreb_ms, traw = kernel_time(e, ktime_steps)
    sps = steps / wall
    info = e.info()
    persistent = e.persist_info()["runs"] > 0   # small single system on the persistent kernel
    traffic, tsrc = traffic_of(name, kernel="persist_kernel" if persistent else "step_kernel")
    out = {
        "value": sps, "unit": "timesteps/s", "ms_per_step": 1e3 * wall / steps,
        "ns_per_day": sps * dt * 0.0864, "steps": steps, "warmup": warmup,
        "rebuilds_per_step": (reb1 - reb0) / steps / max(len(systems) if isinstance(systems, list) else 1, 1),
        "launch": dict(e.get_launch(), list_cap=info["cap"]),
        "step_kernel_ms": kstep_ms,
        "rebuild_chain_ms_per_step": traw["rebuild_ms"] / max(ktime_steps, 1),
        "roofline": {"bound": "valu", "achieved": flops / (kstep_ms * 1e-3) / 1e12, "peak": PEAK_FP32,
                     "unit": "TFLOP/s", "frac": flops / (kstep_ms * 1e-3) / 1e12 / PEAK_FP32,
                     "traffic": traffic, "traffic_source": tsrc, "algorithmic_flop_per_launch": flops,
                     "per": "step of the persistent kernel (one launch runs the whole run)" if persistent
                     else "launch (one step)"},
        "step_roofline_frac": flops / (1e-3 * 1e3 * wall / steps) / 1e12 / PEAK_FP32,
    }
    if long_ms is not None:
        out["long_run"] = {"steps": long_steps, "ms_per_step": long_ms, "value": 1e3 / long_ms,
                           "short_over_long": (1e3 * wall / steps) / long_ms}
    if calibrate:
        nbytes = 64 << 20
        cms = e.copy_bench(nbytes, 5)
        out["hbm_copy_GBps"] = 2 * nbytes / (cms * 1e-3) / 1e9   # read + write of a 64 MiB device copy
        out["rebuild_ms"] = e.rebuild_bench(10)                    # one forced full list rebuild
    if force_check is not None:
        out["force_rel_rms"] = force_check(e, systems)
    if cpu_fn is not None:
        out["cpu_baseline"] = cpu_fn(e, systems)
    e.close()
    return out


def oracle_cpu_steps(s, budget_s, min_steps=2):
    """FP64 oracle trajectory on host cores: steps/s on a bounded sample."""
    import oracle
    t0 = time.perf_counter()
    oracle.md_run(s, 1)
    one = time.perf_counter() - t0
    n = max(min_steps, int(budget_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    oracle.md_run(s, n)
    dt_ = time.perf_counter() - t0
    return n, n / dt_, oracle.num_threads()


def force_error(e, s):
    import oracle
    pos, vel, _ = e.download()
    F, _, _ = e.compute_forces()
    t = s.copy()
    t.pos = pos
    Fo, _, _ = oracle.forces(t)
    return float(np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--quick", action="store_true", help="headline workload only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch", default="", help="fixed launch config for the headline workload, "
                    "e.g. lanes=8,grid=2048,cell_div=2,check_every=16,build_lanes=8 (skips the autotuner)")
    args = ap.parse_args()
    launch = dict((k, int(v)) for k, v in (kv.split("=") for kv in args.launch.split(",") if kv)) or None
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported by this build "
              "(the step path shards by spatial domains; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from md_gpu_amd import engine as E, systems as S
    log("library", E.version())
    want_cpu = not args.no_cpu
    tune = not args.no_autotune

    # ---- headline: 512 batched Ar-Xe replicas, NVT leapfrog Berendsen ----
    reps = S.arxe_replicas()

    def f_batch(p):
        off, tot = 0, 0
        for r in reps:
            tot += count_pairs(p[off:off + r.n], r.box, r.ff.rc)[0]
            off += r.n
        return 36.0 * tot

    res = step_workload("arxe512", reps, f_batch, args.steps, args.warmup, tune, reps[0].dt,
                        (lambda e, s: cpu_batch(e, s, 10.0)) if want_cpu else None,
                        (lambda e, s: batch_force_error(e, s)) if want_cpu else None, launch, calibrate=True,
                        long_steps=0 if args.quick else 500)
    natoms = int(sum(r.n for r in reps))
    line = {
        "metric": METRIC, "value": res["value"], "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "arxe512_batched_nvt_leapfrog_berendsen", "replicas": len(reps), "atoms": natoms,
                   "atoms_min": int(min(r.n for r in reps)), "atoms_max": int(max(r.n for r in reps)),
                   "rc_A": reps[0].ff.rc, "skin_A": reps[0].ff.skin, "dt_fs": reps[0].dt, "T_K": "80..300",
                   "tau_fs": reps[0].tau},
        "ns_per_day": res["ns_per_day"], "force_rel_rms": res.get("force_rel_rms"),
        "roofline": res["roofline"], "cpu_baseline": res.get("cpu_baseline"),
        "step_kernel_ms": res["step_kernel_ms"], "rebuild_chain_ms_per_step": res["rebuild_chain_ms_per_step"],
        "rebuilds_per_step": res["rebuilds_per_step"], "launch": res["launch"],
        "rebuild_ms": res.get("rebuild_ms"), "hbm_copy_GBps": res.get("hbm_copy_GBps"),
        "long_run": res.get("long_run"),
    }
    work = {}
    if not args.quick:
        work = extra_workloads(args, want_cpu, tune)
    line["workloads"] = work
    print(json.dumps(line), flush=True)


def batch_force_error(e, reps, nsample=8):
    """Relative RMS force error against the FP64 oracle over a seeded sample of replicas."""
    import oracle
    pos, _, _ = e.download()
    F, _, _ = e.compute_forces()
    offs = np.concatenate([[0], np.cumsum([r.n for r in reps])])
    rng = np.random.default_rng(1)
    num = den = 0.0
    for k in rng.choice(len(reps), nsample, replace=False):
        t = reps[k].copy()
        t.pos = pos[offs[k]:offs[k + 1]]
        Fo, _, _ = oracle.forces(t)
        num += float(((F[offs[k]:offs[k + 1]] - Fo) ** 2).sum())
        den += float((Fo ** 2).sum())
    return float(np.sqrt(num / den))


def extra_workloads(args, want_cpu, tune):
    from md_gpu_amd import systems as S
    out = {}
    # water 196,608 atoms, NVT velocity Verlet + Nose-Hoover
    wa = S.water()

    def f_water(p):
        n, noo = count_pairs(p, wa.box, wa.ff.rc, wa.type, lambda a, b: (a == 0) & (b == 0))
        return 60.0 * n + 5.0 * noo
    out["water196608_nvt_vv_nh"] = step_workload(
        "water196608_nvt_vv_nh", wa, f_water, 1000, 200, tune, wa.dt,
        (lambda e, s: cpu_small(e, s, 10.0)) if want_cpu else None)
    # molten NaCl 32,768 ions, NVT leapfrog Berendsen
    nacl = S.nacl()

    def cpu_nacl(e, s):
        pos, vel, _ = e.download()
        t = s.copy()
        t.pos, t.vel = pos, vel
        n, sps, cores = oracle_cpu_steps(t, 12.0)
        return {"value": sps, "unit": "timesteps/s", "cores": cores, "kind": "port",
                "sample": f"{n} steps of the same 32,768-ion NaCl state, FP64 oracle (OpenMP, {cpu_info()})"}
    out["nacl32768_wolf_nvt_lf"] = step_workload(
        "nacl32768_wolf_nvt_lf", nacl, lambda p: 65.0 * count_pairs(p, nacl.box, nacl.ff.rc)[0], 5000, 300, tune,
        nacl.dt,
        cpu_nacl if want_cpu else None, (lambda e, s: force_error(e, s)) if want_cpu else None)
    # the same NaCl with Ewald electrostatics: damped real space (alpha 3.09 / rc)
    # plus smooth PME, order 4 on a 96^3 grid (0.94 A)
    nacl_pme = S.with_pme(nacl, spacing=1.0, order=4)

    def cpu_nacl_pme(e, s):
        pos, vel, _ = e.download()
        t = s.copy()
        t.pos, t.vel = pos, vel
        n, sps, cores = oracle_cpu_steps(t, 12.0)
        return {"value": sps, "unit": "timesteps/s", "cores": cores, "kind": "port",
                "sample": f"{n} steps of the same state, FP64 oracle with FP64 SPME (OpenMP, {cpu_info()})"}
    out["nacl32768_pme_nvt_lf"] = step_workload(
        "nacl32768_pme_nvt_lf", nacl_pme, lambda p: 65.0 * count_pairs(p, nacl.box, nacl.ff.rc)[0], 2000, 200,
        tune, nacl.dt, cpu_nacl_pme if want_cpu else None,
        (lambda e, s: force_error(e, s)) if want_cpu else None)
    # rigid water 65,536 molecules with SETTLE and PME (128^3 grid), leapfrog Berendsen 2 fs
    wr = S.with_pme(S.water(rigid=True), spacing=1.0, order=4)

    def f_wr(p):
        n, noo = count_pairs(p, wr.box, wr.ff.rc, wr.type, lambda a, b: (a == 0) & (b == 0))
        return 60.0 * n + 5.0 * noo
    out["water196608_rigid_pme_nvt_lf"] = step_workload(
        "water196608_rigid_pme_nvt_lf", wr, f_wr, 500, 100, tune, wr.dt,
        (lambda e, s: cpu_small(e, s, 10.0)) if want_cpu else None)
    # argon 864, NVE velocity Verlet (the 1000-step case, timed over 5000 steps: a
    # step takes microseconds)
    ar = S.argon()
    out["argon864_nve_vv"] = step_workload(
        "argon864_nve_vv", ar, lambda p: 36.0 * count_pairs(p, ar.box, ar.ff.rc)[0], 5000, 500, tune, ar.dt,
        (lambda e, s: cpu_small(e, s, 5.0)) if want_cpu else None)
    out["nacl4096x64_force_matching"] = fm_workload(want_cpu)
    return out


def cpu_small(e, s, budget):
    pos, vel, _ = e.download()
    t = s.copy()
    t.pos, t.vel = pos, vel
    n, sps, cores = oracle_cpu_steps(t, budget)
    return {"value": sps, "unit": "timesteps/s", "cores": cores, "kind": "port",
            "sample": f"{n} steps of the same state, FP64 oracle"}


def cpu_batch(e, reps, budget):
    """Oracle on a seeded subset of replicas; steps/s of the whole batch
    extrapolated by atom count (the oracle cost is linear in atoms)."""
    import oracle
    rng = np.random.default_rng(0)
    pick = rng.choice(len(reps), 16, replace=False)
    t0 = time.perf_counter()
    atoms = 0
    nst = 5
    for k in pick:
        oracle.md_run(reps[k], nst)
        atoms += reps[k].n
    el = time.perf_counter() - t0
    total = sum(r.n for r in reps)
    sps = nst / (el * total / atoms)
    return {"value": sps, "unit": "timesteps/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{nst} steps of 16 seeded replicas ({atoms} atoms), scaled to {total} atoms"}


def fm_workload(want_cpu):
    """Force matching on NaCl 4096 ions x 64 frames: loss + gradient over 7 parameters."""
    from md_gpu_amd import engine as E
    from md_gpu_amd.fit import nacl_problem
    log("fm: generating 64 frames by GPU MD and the perturbed model's target forces")
    e, fm, base, frames, tgt, _ = nacl_problem(E.Engine, nframes=64)
    theta = fm.theta0()
    fm.loss_grad(theta)
    n = 50
    t0 = time.perf_counter()
    for _ in range(n):
        loss, g = fm.loss_grad(theta)
    el = time.perf_counter() - t0
    # algorithmic flops per eval: (65 forward + 45 adjoint) per ordered... per unordered pair x 2 directions
    P = np.mean([count_pairs(f, base.box, base.ff.rc)[0] for f in frames[:4]])
    flops = 64 * P * (65 + 45)
    # the per-tile force-matching kernel alone, HIP events on the engine stream
    e.timing(True)
    for _ in range(20):
        fm.loss_grad(theta)
    t = e.timing_get()
    e.timing(False)
    kms = t["step_ms"] / max(t["step_launches"], 1)
    traffic, tsrc = traffic_of("nacl4096x64", kernel="fm_kernel")
    out = {"value": n / el, "unit": "force-grad evals/s", "ms_per_eval": 1e3 * el / n, "loss": loss,
           "fm_kernel_ms": kms,
           "roofline": {"bound": "valu", "achieved": flops / (kms * 1e-3) / 1e12, "peak": PEAK_FP32, "unit": "TFLOP/s",
                        "frac": flops / (kms * 1e-3) / 1e12 / PEAK_FP32, "traffic": traffic,
                        "traffic_source": tsrc, "algorithmic_flop_per_launch": flops},
           "eval_roofline_frac": flops / (el / n) / 1e12 / PEAK_FP32}
    if want_cpu:
        out["cpu_baseline"] = cpu_fd_baseline(fm, base, frames, tgt, theta, nframes=4)
    e.close()
    return out


def cpu_fd_baseline(fm, base, frames, tgt, theta, nframes=4, rel_h=1e-4):
    """FP64 oracle central finite differences over the 7 parameters (14 loss
    evaluations + 1), on a subset of frames, scaled to 64 frames."""
    import oracle
    fr = frames[:nframes]
    tg = tgt[:nframes].astype(np.float64)

    def loss(th):
        lj1, lj2, q = fm.tables(th)
        return oracle.fm_loss_grad(fr, base.type, base.charge, base.box, base.ff, q, tg, lj1, lj2)[0]
    t0 = time.perf_counter()
    loss(theta)
    for k in range(len(theta)):
        h = rel_h * abs(theta[k])
        tp = theta.copy(); tp[k] += h
        tm = theta.copy(); tm[k] -= h
        loss(tp); loss(tm)
    el = time.perf_counter() - t0
    scale = frames.shape[0] / nframes
    return {"value": 1.0 / (el * scale), "unit": "force-grad evals/s", "cores": oracle.num_threads(),
            "kind": "port", "sample": f"central FD (15 FP64 loss evals) on {nframes} of 64 frames, scaled"}


if __name__ == "__main__":
    main()
