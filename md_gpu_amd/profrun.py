"""Short fixed-launch run of one benchmark workload for hardware-counter
collection and kernel traces: molten NaCl (32,768 ions, the default), argon
(864 atoms, NVE velocity Verlet), water (196,608 atoms, NVT velocity Verlet),
the same NaCl and rigid water (SETTLE) with smooth PME (order 4, about 1 A
grid: 96^3 and 128^3), the 512 batched Ar-Xe replicas, force matching (NaCl 4,096 ions x 64 frames,
--steps loss-and-gradient evaluations) or a plain 64 MiB device copy (the
calibration of the HBM byte counters).

    python -m md_gpu_amd.profrun [--system nacl|argon|water|arxe|nacl_pme|water_rigid_pme|fm|copy]
                                 [--launch k=v,...] [--steps N] [--timed]

Counter collection serialises every dispatch, so this driver keeps the run
short, needs no autotuning and reports (rather than raises) a failed run.
With --timed every step is launched directly with HIP events around it (the
same timing path bench.py uses for kernel durations).

The JSON line ends with a checksum of the final state (FP64 sums of the
downloaded positions and velocities, and their hex digest) and the device
list-invariant counters.  The engine is deterministic, so a run under a
profiler is trustworthy only if its checksum equals that of the same command
run without one.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import sys
import time

import numpy as np


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--launch", default="lanes=8,cell_div=2,check_every=8,refresh=1,defer=1,submit=1")
    ap.add_argument("--system", default="nacl", choices=["nacl", "argon", "water", "arxe", "nacl_pme", "water_rigid_pme", "fm", "copy"])
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--timed", action="store_true")
    args = ap.parse_args()
    from md_gpu_amd import engine as E, systems as S
    cfg = dict((k, int(v)) for k, v in (kv.split("=") for kv in args.launch.split(",") if kv))
    if args.system == "copy":
        e = E.Engine.from_system(S.argon(ncell=4))
        ms = e.copy_bench(64 << 20, 6)
        print(json.dumps(dict(ok=True, system="copy", bytes=64 << 20, reps=6, ms=ms)), flush=True)
        e.close()
        return 0
    if args.system == "fm":
        return fm_main(args, cfg)
    build = {"nacl": S.nacl, "argon": S.argon, "water": S.water, "arxe": S.arxe_replicas,
             "nacl_pme": lambda: S.with_pme(S.nacl()),
             "water_rigid_pme": lambda: S.with_pme(S.water(rigid=True))}[args.system]
    e = E.Engine.from_system(build())
    e.set_launch(**cfg)
    try:
        if args.warmup:
            e.run(args.warmup, stats=False)
        if args.timed:
            e.timing(True)
        reb0 = e.info()["rebuilds"]
        t0 = time.perf_counter()
        e.run(args.steps, stats=False)
        wall = time.perf_counter() - t0
        out = dict(ok=True, system=args.system, launch=e.get_launch(), steps=args.steps,
                   rebuilds=e.info()["rebuilds"], run_rebuilds=e.info()["rebuilds"] - reb0,
                   wall_ms_per_step=1e3 * wall / args.steps)
        if args.timed:
            out["timing"] = e.timing_get()
    except E.MDError as x:
        out = dict(ok=False, error=str(x))
    try:
        pos, vel, _ = e.download()
        out["checksum"] = dict(pos=float(pos.astype("f8").sum()), vel=float(vel.astype("f8").sum()),
                               digest=hashlib.sha1(pos.tobytes() + vel.tobytes()).hexdigest()[:16])
        out["invariants"] = e.check_lists()
    except E.MDError as x:
        out["checksum"] = dict(error=str(x))
    print(json.dumps(out), flush=True)
    e.close()
    return 0


def fm_main(args, cfg) -> int:
    """Force matching: --steps evaluations of loss and gradient at the base
    parameters; the checksum is the loss and gradient themselves."""
    from md_gpu_amd import engine as E
    from md_gpu_amd.fit import nacl_problem
    e, fm, base, frames, tgt, _ = nacl_problem(E.Engine, nframes=64)
    e.set_launch(**cfg)
    th = fm.theta0()
    if args.warmup:
        fm.loss_grad(th)
    if args.timed:
        e.timing(True)
    for _ in range(args.steps):
        loss, g = fm.loss_grad(th)
    out = dict(ok=True, system="fm", launch=e.get_launch(), evals=args.steps, loss=loss, grad=[float(x) for x in g])
    out["checksum"] = dict(digest=hashlib.sha1(np.float64([loss, *g]).tobytes()).hexdigest()[:16])
    print(json.dumps(out), flush=True)
    e.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
