"""Throughput of the per-timestep finite-volume update on one MI355X.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full RK2 timestep (getmindt + 2 substeps, timestep.c:7-54) of
every real cell of the workload, with the state resident in HBM; the value is
zone-updates/s = real cells x K / wall time of the K graph-launched steps
(HIP events on the library's stream, plus a device sync on both sides).
The headline workload is the 3D ideal-MHD disk (HLLD, constrained
transport, Nr=128 x Nz=64, 5,512,000 real cells); the 2D hydro disk, the
3D GRMHD torus (Kerr-Schild a=0.9, 11.7 M cells), the 48-disk batch and the
standalone GRMHD cons2prim are reported beside it.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import hashlib
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
# compulsory HBM bytes of one zone-substep of a perfectly fused update:
# hydro (NUM_Q=5, no CT): read prim, cons, RKcons (15 doubles) + piph, dphi
# (2); write cons, prim (10) + piph, dphi (2) = 29 doubles = 232 B.
# MHD with 3D CT (NUM_Q=8, NUM_FACES=5): read prim, cons, RKcons (24) + Phi,
# RK_Phi (10) + piph, dphi (2); write cons, prim (16) + Phi (5) + piph, dphi
# (2) = 59 doubles = 472 B.
# GRMHD with 3D CT moves the same 59 doubles as MHD.
BYTES_PER_ZONE_SUBSTEP = {"euler": 232, "mhd": 472, "grmhd": 472}

WORKLOADS = {
    # 3D cylindrical ideal MHD disk, HLLD + constrained transport, Nr=128 in
    # r in [1, 4], Nz=64 in z in [-0.5, 0.5] periodic (5,512,000 real cells)
    "mhd3d_nr128_nz64": ("c2_mhd3d.par", {}, "mhd", "mhd_sun"),
    # 2D cylindrical Newtonian hydro Keplerian disk, Nr=256, r in [0.5, 3]
    "kepler2d_nr256": ("c1_kepler2d.par", {}, "euler", "euler_sun"),
    # 3D GRMHD torus in Kerr-Schild (a = 0.9), HLL + constrained transport,
    # Nr=192 log-spaced in [1.3, 60], Nz=192 in [-30, 30], fixed boundaries
    # (11,685,888 real cells)
    "grmhd_torus_nr192_nz192": ("c5_grmhd_torus.par", {}, "grmhd", "grmhd_torus"),
}
# bounded CPU sample of each workload for the reference build: the same grid
# spacing and physics on a slab of the domain, a few steps
CPU_SAMPLES = {
    "mhd3d_nr128_nz64": ({"Num_Z": 8, "Z_Min": -0.0625, "Z_Max": 0.0625}, 3,
                         "8 of the 64 z-layers (same dz, periodic), 689,000 real cells, 3 RK2 steps"),
    "kepler2d_nr256": ({}, 20, "full 288,917-cell grid, 20 RK2 steps"),
    # the reference's metric is Schwarzschild (schw_ks_rpz.c): spin 0
    "grmhd_torus_nr192_nz192": ({"Num_Z": 8, "Z_Min": -1.25, "Z_Max": 1.25, "Metric_Par3": 0.0}, 1,
                                "8 midplane z-layers of the 192 (same dz, fixed z boundaries), spin a = 0 "
                                "(the reference's metric is Schwarzschild), 486,912 real cells, 1 RK2 step"),
}


def kernel_sources_digest() -> str:
    """SHA-256 over the device and C-ABI sources (disco_amd/csrc, include):
    the build a set of rocprofv3 counters belongs to."""
    h = hashlib.sha256()
    for d in ("disco_amd/csrc", "include"):
        for name in sorted(os.listdir(os.path.join(ROOT, d))):
            if name.endswith((".hip", ".h", ".cpp")):
                h.update(name.encode())
                with open(os.path.join(ROOT, d, name), "rb") as f:
                    h.update(f.read())
    return h.hexdigest()


def par_text(base: str, **kv) -> str:
    text = open(os.path.join(ROOT, "configs", base)).read()
    for k, v in kv.items():
        text = re.sub(rf"^{k}\s.*$", f"{k} {v}", text, flags=re.M)
    return text


def cpu_baseline(workload: str) -> dict | None:
    """The reference CPU build (oracle/_ref, compiled from the reference's own
    sources) on a bounded sample of the same workload, serial (the reference
    has no threads), timed around its step loop."""
    base, over, _, variant = WORKLOADS[workload]
    samp, nsteps, what = CPU_SAMPLES[workload]
    exe = os.path.join(ROOT, "oracle", "_ref", f"disco_ref_{variant}")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "in.par"), "w").write(par_text(base, **{**over, **samp}))
        out = subprocess.run([exe, str(nsteps), "-"], cwd=d, capture_output=True, text=True, timeout=300).stdout
    m = re.search(r"steps=(\d+) cells=(\d+) seconds=([0-9.eE+-]+)", out)
    if not m:
        return None
    steps, cells, secs = int(m.group(1)), int(m.group(2)), float(m.group(3))
    return {"value": cells * steps / secs, "unit": "zone-updates/s", "cores": 1, "kind": "reference",
            "sample": f"{what}: {cells} cells x {steps} steps in {secs:.2f} s (reference C build, gcc -O3, serial)"}


def bench_fv(workload: str, steps: int, warmup: int, tune: bool = True) -> dict:
    """Zone-updates/s of one workload, its per-kernel times and the roofline
    of one RK substep (all kernels of the substep) against HBM.  The
    library's autotuner first picks the launch shape (annulus tile, threads
    per block) on the run itself, outside the timed region."""
    import disco_amd as D

    base, over, hydro, _ = WORKLOADS[workload]
    dom = D.Domain(D.parse_par(par_text(base, **over)))
    nreal = sum(i.nreal for i in dom.info)
    tuned = dom.autotune(2) if tune else None
    dom.step(warmup)
    dom.sync()
    t0 = time.perf_counter()
    ms = dom.time_steps(steps)
    dom.sync()
    wall = time.perf_counter() - t0
    value = nreal * steps / (ms * 1e-3)
    # per-kernel durations (HIP events around eager launches on the same stream)
    kt = dom.kernel_times(5)
    dom.close()
    per_step = {"k_dt_final", "k_planets", "k_diag", "k_diag_time"}
    sub = [k for k in kt if k not in per_step and not k.startswith(("k_flip", "k_mindt", "k_dt_init"))]
    # average substep: every substep kernel once, the flip kernels every other substep
    substep_ms = sum(kt[k][0] for k in sub) + 0.5 * sum(kt[k][0] for k in kt if k.startswith("k_flip"))
    dominant = max(kt, key=lambda k: kt[k][0])
    bpz = BYTES_PER_ZONE_SUBSTEP[hydro]
    achieved = bpz * nreal / (substep_ms * 1e-3) / 1e9
    # measured HBM bytes of one substep, from the rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs (calibrated on a copy kernel; the summary the file names
    # holds the commands).  Reported only while the kernel sources are the
    # ones those counters were taken on: any change to them drops the figure.
    traffic = traffic_src = None
    tf = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
        if t.get("sources_sha256") == kernel_sources_digest():
            traffic = t.get("bytes_per_substep")
            traffic_src = f"{t.get('source')} (counters of this build's sources, commit {t.get('commit', '?')})"
        else:
            traffic_src = f"stale: {t.get('source')} holds counters of other kernel sources (commit {t.get('commit', '?')})"
    return {
        "value": value, "ms_per_step": ms / steps, "wall_s_timed": wall, "real_cells": nreal,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": f"one RK substep ({' + '.join(sub)}; flip kernels every other substep)",
                     "bytes_per_zone_substep": bpz, "substep_ms": substep_ms, "dominant_kernel": dominant},
        "kernels_ms": {k: v[0] for k, v in kt.items()},
        "autotune": tuned,
    }


def bench_grmhd_learned(steps: int, warmup: int, train_steps: int = 1500) -> dict:
    """The GRMHD torus with the learned cons2prim start in the loop: the guess
    network is trained on the GPU on (cons, prim) pairs of the run's own cells
    after its first two steps, then the same run is timed with the plain start
    (previous prims, as the reference) and with the learned one; Newton
    iterations per solve and the fallback rate are counted over one step of
    each."""
    import disco_amd as D

    base, over, _, _ = WORKLOADS["grmhd_torus_nr192_nz192"]
    p = D.parse_par(par_text(base, **over))
    dom = D.Domain(p)
    nreal = dom.info[0].nreal
    dom.step(2)
    x, dV, real, st = dom.cell_positions()
    cons = st["cons"][:, real] / dV
    net = D.C2P(capacity=cons.shape[1], gamma=p.Adiabatic_Index, M=p.metricPar2, a=p.metricPar3)
    net.upload(cons, x)
    t0 = time.perf_counter()
    loss = net.train_resident(seed=7, steps=train_steps, batch=4096, lr=2e-3)
    train_s = time.perf_counter() - t0
    w, nrm = net.get_mlp()
    net.close()
    del cons, x, dV, st
    out = {"train_pairs": int(real.sum()), "train_steps": train_steps, "train_seconds": train_s,
           "train_loss_first_last": [float(loss[:20].mean()), float(loss[-20:].mean())]}
    for mode in ("plain", "learned"):
        dom.set_c2p_net(w, nrm) if mode == "learned" else dom.set_c2p_net(None)
        dom.c2p_count(False)
        dom.step(warmup)
        dom.sync()
        ms = dom.time_steps(steps)
        kt = dom.kernel_times(2)
        dom.c2p_count(True)
        dom.c2p_stats(reset=True)
        dom.step(1)
        c = dom.c2p_stats(reset=True)
        upd = "k_update_b_learned" if mode == "learned" else "k_update_b"
        out[mode] = {"value": nreal * steps / (ms * 1e-3), "ms_per_step": ms / steps, "update_b_ms": kt[upd][0],
                     "iterations_per_solve": c["iterations"] / c["solves"],
                     "fallback_rate": c["fallbacks"] / c["solves"], "unconverged": c["unconverged"]}
    out["speedup_learned_vs_plain"] = out["plain"]["ms_per_step"] / out["learned"]["ms_per_step"]
    out["update_b_speedup"] = out["plain"]["update_b_ms"] / out["learned"]["update_b_ms"]
    dom.close()
    return out


def bench_batch(steps: int, warmup: int) -> dict:
    """48 independent binary disks of uneven size (seeded Nr in [32, 256])
    advanced together, against the same 48 runs advanced one after another
    (each its own domain, synchronised before the next starts)."""
    import disco_amd as D
    from disco_amd.workloads import binary_batch

    runs = binary_batch(48)
    dom = D.Domain(runs)
    cells = sum(i.nreal for i in dom.info)
    tuned = dom.autotune(3)  # threads per block of the batched launches
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
    per_step = {"k_dt_final", "k_planets", "k_diag", "k_diag_time"}
    substep_ms = sum(v[0] for k, v in kt.items() if k not in per_step and not k.startswith(("k_mindt", "k_dt_init")))
    achieved = BYTES_PER_ZONE_SUBSTEP["euler"] * cells / (substep_ms * 1e-3) / 1e9
    return {"workload": "binary_batch_x48", "runs": 48, "real_cells": cells, "steps": steps,
            "value": batched, "unit": "zone-updates/s", "ms_per_step": ms / steps,
            "sequential_value": sequential, "speedup_vs_sequential": batched / sequential,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_zone_substep": BYTES_PER_ZONE_SUBSTEP["euler"],
                         "substep_ms": substep_ms},
            "kernels_ms": {k: v[0] for k, v in kt.items()}, "autotune": tuned}


# FP64 operations of the plain cons2prim per state, counted by hand from
# disco_amd/csrc/phys_grmhd.h (divide and sqrt count as 1): Kerr-Schild metric
# and frame 110, invariants 72, cold start 5, recovery 58 => 245 per state, and
# 61 per Newton iteration.
C2P_FLOP_STATE, C2P_FLOP_ITER = 245, 61
MLP_FLOP_SAMPLE = 2 * (8 * 64 + 3 * 64 * 64 + 64 * 2)  # 25,856 flops (multiply-add = 2)
FP64_PEAK_TFLOPS, BF16_PEAK_TFLOPS = 78.6, 2500.0


def bench_c2p(nstates: int, train_steps: int) -> dict:
    """Standalone GRMHD cons2prim on 2^24 seeded Kerr-Schild (a = 0.9) states:
    plain cold start vs learned start (BF16-MFMA MLP + FP64 polish), both with
    the reference's Nextra = 10 extra iterations and with Nextra = 0."""
    import numpy as np

    import disco_amd as D
    from disco_amd.workloads import c4_params

    P = c4_params()
    h = D.C2P(nstates, P["gamma"], P["M"], P["a"], rmin=P["rmin"], rmax=P["rmax"], theta_cut=P["theta_cut"])
    t0 = time.perf_counter()
    loss = h.train(seed=P["seed"] + 1, ntrain=1 << 20, steps=train_steps, batch=8192, lr=2e-3)
    train_s = time.perf_counter() - t0
    h.generate(P["seed"], 0, nstates)
    out = {"workload": "grmhd_c2p_kerr_a0.9_2^24", "states": nstates, "train_steps": train_steps,
           "train_seconds": train_s, "train_loss_first_last": [float(loss[0]), float(loss[-5:].mean())]}
    for ne in (0, 10):
        h.solve(D.PLAIN, nextra=ne)  # warm
        tp = min(h.solve(D.PLAIN, nextra=ne) for _ in range(3))
        a = h.download()
        ep = h.errors(P["seed"], 0)
        h.solve(D.LEARNED, nextra=ne)
        tl = [h.solve(D.LEARNED, nextra=ne, split=True) for _ in range(3)]
        tl, tg = min(t[0] for t in tl), min(t[1] for t in tl)
        b = h.download()
        el = h.errors(P["seed"], 0)
        q, r = a["prim"], b["prim"]
        rel = np.abs(r - q) / np.maximum(np.abs(q), 1e-300)
        rel[2:5] = np.abs(r[2:5] - q[2:5]) / np.maximum(np.abs(q[2:5]).max(0), 1e-300)
        it_p, it_l = float(a["iters"].mean()), float(b["iters"].mean())
        flops = nstates * (C2P_FLOP_STATE + C2P_FLOP_ITER * it_p)
        out[f"nextra{ne}"] = {
            "plain_states_per_s": nstates / (tp * 1e-3), "learned_states_per_s": nstates / (tl * 1e-3),
            "speedup": tp / tl, "plain_ms": tp, "learned_ms": tl, "learned_network_ms": tg,
            "plain_iters_mean": it_p, "learned_iters_mean": it_l,
            "fallback_rate": float((b["flags"] % 2 == 1).mean()),
            "unconverged_plain": int((a["flags"] >= 2).sum()), "unconverged_learned": int((b["flags"] >= 2).sum()),
            "max_rel_diff_learned_vs_plain": float(np.nanmax(rel[[0, 2, 3, 4]])),
            "max_rel_diff_P_learned_vs_plain": float(np.nanmax(rel[1])),
            "err_vs_truth_plain_p50_p99_max": [float(np.percentile(ep, 50)), float(np.percentile(ep, 99)),
                                               float(ep.max())],
            "err_vs_truth_learned_p50_p99_max": [float(np.percentile(el, 50)), float(np.percentile(el, 99)),
                                                 float(el.max())],
            "plain_fp64_tflops": flops / (tp * 1e-3) / 1e12,
            "network_bf16_tflops": nstates * MLP_FLOP_SAMPLE / (tg * 1e-3) / 1e12,
        }
    h.close()
    r0 = out["nextra0"]
    out["roofline"] = {"bound": "fp64", "kernel": "k_c2p_plain (Nextra=0)", "achieved": r0["plain_fp64_tflops"],
                       "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": r0["plain_fp64_tflops"] / FP64_PEAK_TFLOPS,
                       "flop_per_state": f"{C2P_FLOP_STATE} + {C2P_FLOP_ITER} x iterations"}
    out["mfma_roofline"] = {"bound": "mfma", "kernel": "k_mlp_guess", "achieved": r0["network_bf16_tflops"],
                            "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": r0["network_bf16_tflops"] / BF16_PEAK_TFLOPS, "flop_per_sample": MLP_FLOP_SAMPLE}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mhd3d_nr128_nz64", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="only the headline workload")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-c2p", action="store_true")
    ap.add_argument("--c2p-states", type=int, default=1 << 24)
    ap.add_argument("--c2p-train-steps", type=int, default=2000)
    ap.add_argument("--progress", action="store_true", help="a line on stderr after each workload")
    ap.add_argument("--only", choices=["batch", "c2p"], default=None,
                    help="run only the 48-disk batch or the standalone cons2prim leg (profiling)")
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

    def note(msg: str) -> None:  # progress on stderr; the JSON line stays alone on stdout
        if args.progress:
            print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.only == "batch":
        print(json.dumps({"batch": bench_batch(args.steps, args.warmup)}))
        return 0
    if args.only == "c2p":
        print(json.dumps({"c2p": bench_c2p(args.c2p_states, args.c2p_train_steps)}))
        return 0
    wl = args.workload
    r = bench_fv(wl, args.steps, args.warmup)
    note(f"{wl}: {r['value']:.4g} zone-updates/s")
    hydro = WORKLOADS[wl][2]
    cfg = {"workload": wl, "real_cells": r["real_cells"], "integrator": "RK2", "plm": 1.5,
           "reference_convention_megazones_per_s": 2 * r["value"] / 1e6}
    if hydro == "grmhd":
        cfg.update({"physics": "GRMHD (grmhd.c), Kerr-Schild a=0.9", "riemann": "HLL", "ct": "CT_MODE 2 (3D)",
                    "Nr": 192, "Nz": 192, "z": "fixed", "mesh_motion": "Keplerian, rigid per annulus"})
    elif hydro == "mhd":
        cfg.update({"physics": "ideal MHD (mhd.c)", "riemann": "HLLD", "ct": "CT_MODE 2 (3D)", "Nr": 128, "Nz": 64,
                    "z": "periodic", "mesh_motion": "Keplerian, rigid per annulus"})
    else:
        cfg.update({"physics": "Newtonian hydro (euler.c)", "riemann": "HLLC"})
    line = {
        "metric": "zone-updates/s (FP64) 1 GPU",
        "value": r["value"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded disk, identical to the reference's initial data)",
        "config": cfg,
        "roofline": r["roofline"],
        "kernels_ms": r["kernels_ms"],
        "autotune": r["autotune"],
        "wall_s_timed": r["wall_s_timed"],
    }
    if not args.no_extra and wl != "kepler2d_nr256":
        k = bench_fv("kepler2d_nr256", 200, 20)
        k["workload"] = "kepler2d_nr256"
        if not args.no_cpu:
            k["cpu_baseline"] = cpu_baseline("kepler2d_nr256")
        line["kepler2d"] = k
        note(f"kepler2d_nr256: {k['value']:.4g} zone-updates/s")
    if not args.no_extra and wl != "grmhd_torus_nr192_nz192":
        g = bench_fv("grmhd_torus_nr192_nz192", 10, 2)
        g["workload"] = "grmhd_torus_nr192_nz192"
        note(f"grmhd_torus_nr192_nz192: {g['value']:.4g} zone-updates/s")
        g["learned_c2p"] = bench_grmhd_learned(10, 2)
        note("grmhd torus with the learned cons2prim start done")
        if not args.no_cpu:
            g["cpu_baseline"] = cpu_baseline("grmhd_torus_nr192_nz192")
        line["grmhd_torus"] = g
    if not args.no_batch and not args.no_extra:
        line["batch"] = bench_batch(args.steps, args.warmup)
        note(f"batch: {line['batch']['value']:.4g} zone-updates/s")
    if not args.no_c2p and not args.no_extra:
        line["c2p"] = bench_c2p(args.c2p_states, args.c2p_train_steps)
        note("cons2prim done")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
