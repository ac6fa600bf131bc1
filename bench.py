"""Throughput of the per-timestep finite-volume update on one MI355X.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full RK2 timestep (getmindt + 2 substeps, timestep.c:7-54) of
every real cell of the workload, with the state resident in HBM; the value is
zone-updates/s = real cells x K / wall time of the K graph-launched steps
(HIP events on the library's stream, plus a device sync on both sides).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
# compulsory HBM bytes of one zone-substep of the 2D/3D hydro update (NUM_Q=5,
# no CT): read prim, cons, RKcons (15 doubles) + piph, dphi (2); write cons,
# prim (10) + piph, dphi (2) = 29 doubles
BYTES_PER_ZONE_SUBSTEP_HYDRO = 232

WORKLOADS = {
    # 2D cylindrical Newtonian hydro Keplerian disk, Nr=256, r in [0.5, 3]
    "kepler2d_nr256": ("c1_kepler2d.par", {}),
}


def par_text(base: str, **kv) -> str:
    text = open(os.path.join(ROOT, "configs", base)).read()
    for k, v in kv.items():
        text = re.sub(rf"^{k}\s.*$", f"{k} {v}", text, flags=re.M)
    return text


def cpu_baseline(workload: str) -> dict | None:
    """The reference CPU build (oracle/_ref, compiled from the reference's own
    sources) on a bounded sample of the same workload: the full grid for a
    fixed number of steps, serial (the reference has no threads)."""
    base, over = WORKLOADS[workload]
    exe = os.path.join(ROOT, "oracle", "_ref", "disco_ref_euler_sun")
    if not os.path.exists(exe):
        return None
    nsteps = 20
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "in.par"), "w").write(par_text(base, **over))
        out = subprocess.run([exe, str(nsteps), "-"], cwd=d, capture_output=True, text=True, timeout=300).stdout
    m = re.search(r"steps=(\d+) cells=(\d+) seconds=([0-9.eE+-]+)", out)
    if not m:
        return None
    steps, cells, secs = int(m.group(1)), int(m.group(2)), float(m.group(3))
    return {"value": cells * steps / secs, "unit": "zone-updates/s", "cores": 1, "kind": "reference",
            "sample": f"full {cells}-cell grid, {steps} RK2 steps, {secs:.2f} s (reference C build, gcc -O3, serial)"}


def bench_batch(steps: int, warmup: int) -> dict:
    """48 independent binary disks of uneven size (seeded Nr in [32, 256])
    advanced together, against the same 48 runs advanced one after another
    (each its own domain, synchronised before the next starts)."""
    import disco_amd as D
    from disco_amd.workloads import binary_batch

    runs = binary_batch(48)
    dom = D.Domain(runs)
    cells = sum(i.nreal for i in dom.info)
    dom.step(warmup)
    dom.sync()
    ms = dom.time_steps(steps)
    batched = cells * steps / (ms * 1e-3)
    kt = dom.kernel_times(5)
    dom.close()
    singles = [D.Domain(p) for p in runs]
    for d in singles:
        d.step(warmup)
        d.sync()
    t0 = time.perf_counter()
    for d in singles:
        d.step(steps)
        d.sync()
    seq_s = time.perf_counter() - t0
    for d in singles:
        d.close()
    sequential = cells * steps / seq_s
    return {"workload": "binary_batch_x48", "runs": 48, "real_cells": cells, "steps": steps,
            "value": batched, "unit": "zone-updates/s", "ms_per_step": ms / steps,
            "sequential_value": sequential, "speedup_vs_sequential": batched / sequential,
            "kernels_ms": {k: v[0] for k, v in kt.items()}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="kepler2d_nr256", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_bytes:
        import torch

        n = args.copy_bytes // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 8 * n, "launches": 5}))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: one GPU owns the whole domain "
                                   "(no sharding, no collectives); run with --gpus 1"}))
        return 2

    import disco_amd as D

    base, over = WORKLOADS[args.workload]
    dom = D.Domain(D.parse_par(par_text(base, **over)))
    nreal = sum(i.nreal for i in dom.info)
    dom.step(args.warmup)
    dom.sync()
    t0 = time.perf_counter()
    ms = dom.time_steps(args.steps)
    dom.sync()
    wall = time.perf_counter() - t0
    value = nreal * args.steps / (ms * 1e-3)

    # per-kernel durations (HIP events around eager launches on the same stream)
    kt = dom.kernel_times(10)
    sub = ["k_stage_a", "k_plm_r", "k_flux", "k_update"]
    substep_ms = sum(kt[k][0] for k in sub if k in kt)
    dominant = max(sub, key=lambda k: kt.get(k, (0, 0))[0])
    achieved = BYTES_PER_ZONE_SUBSTEP_HYDRO * nreal / (substep_ms * 1e-3) / 1e9
    line = {
        "metric": "zone-updates/s (FP64) 1 GPU",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Keplerian disk, identical to the reference's initial data)",
        "config": {"workload": args.workload, "real_cells": nreal, "riemann": "HLLC", "plm": 1.5,
                   "integrator": "RK2", "reference_convention_megazones_per_s": 2 * value / 1e6},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "one RK substep (k_stage_a + k_plm_r + k_flux + k_update)",
                     "bytes_per_zone_substep": BYTES_PER_ZONE_SUBSTEP_HYDRO,
                     "substep_ms": substep_ms, "dominant_kernel": dominant},
        "kernels_ms": {k: v[0] for k, v in kt.items()},
        "wall_s_timed": wall,
    }
    dom.close()
    if not args.no_batch:
        line["batch"] = bench_batch(args.steps, args.warmup)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
