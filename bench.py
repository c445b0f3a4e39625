#!/usr/bin/env python3
"""Throughput of the MD step path on one MI355X.

Workload: 1,000,000-atom argon, simple-cubic 100^3 at lattice edge 3.338339 nm
(box 333.8339 nm), seeded Maxwell-Boltzmann at 273.15 K, Lennard-Jones
(sigma 0.3418, eps 1.712, r_cut 2.5 sigma) through a Verlet-skin list, FP32
pair forces, FP64 state and reductions, Berendsen thermostat (tau 10, 300 K)
and barostat (beta 1, tau 5, 1.01325), dt 0.002.  A step is one velocity-
Verlet step of the whole system; the state stays resident in HBM.

Prints one JSON line.  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s at 1M-atom LJ NPT on 1 GPU vs FP64 plain; force RMSE vs FP64 oracle"
N_SIDE = 100
A_LAT = 3.338339
MASS = 66.335
DT = 0.002
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def initial_state(M, n_side):
    n = n_side ** 3
    pos = M.lattice("u", [n_side] * 3, A_LAT)
    vel = M.mb_velocities(n, 273.15, MASS, 3)
    box = np.full(3, n_side * A_LAT)
    return n, pos, vel, box


def controls(M):
    th = M.thermostat("berendsen", tau=10.0, target_t=300.0)
    ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=1.01325)
    return th, ba


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n-side", type=int, default=N_SIDE)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--skin", type=float, default=None)
    ap.add_argument("--block-size", type=int, default=256)
    ap.add_argument("--cells-per-atom", type=float, default=2.0)
    ap.add_argument("--tiles-per-block", type=int, default=0)
    ap.add_argument("--graph-chunk", type=int, default=32, help="launches per captured graph (1: none)")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-launch device timestamps in the timed region (no roofline)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tune-steps", type=int, default=300)
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "k_step_1M_npt_traffic.txt"),
                    help="rocprof_summary traffic file holding bytes_per_launch of the step kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(this engine runs one system or replica batch per GPU)", file=sys.stderr)
        sys.exit(2)

    import moldyn_amd as M

    n, pos, vel, box = initial_state(M, args.n_side)
    log(f"[bench] state ready: {n} atoms, box {box[0]:.4f}")
    opts = {"skin": args.skin if args.skin is not None else 0.3 * 0.3418,
            "block_size": args.block_size, "cells_per_atom": args.cells_per_atom,
            "tiles_per_block": args.tiles_per_block}
    if not args.no_tune:
        from moldyn_amd.autotune import autotune
        # tune on the state half-way through the timed run: the barostat
        # compresses the gas several-fold over the run, which lengthens the
        # lists and shortens the rebuild interval
        pre = args.warmup + args.steps // 2
        e0 = M.Engine([n], [MASS], box, force_mode="fp32_list", skin=0.5, cells_per_atom=4.0)
        e0.upload(pos, vel, box)
        e0.update_force()
        th0, ba0 = controls(M)
        e0.step(DT, pre, th0, ba0)
        mid = e0.download(("x", "v", "box"))
        e0.close()
        log(f"[bench] tuning at step {pre}: box {mid['box'][0]:.4f}")
        opts, _ = autotune([n], [MASS], mid["box"], mid["x"], mid["v"], DT, lambda: controls(M),
                           steps=args.tune_steps, warmup=20, log=log)
        log(f"[bench] autotuned options: {opts}")
    e = M.Engine([n], [MASS], box, force_mode="fp32_list", graph_chunk=args.graph_chunk, **opts)
    e.upload(pos, vel, box)
    e.update_force()
    th, ba = controls(M)
    e.step(DT, args.warmup, th, ba)
    log(f"[bench] warmup {args.warmup} steps done")

    # per-launch device timestamps of every kernel in the timed region (the
    # graphs keep running; each launch is timed from its first block's start
    # to its last block's end on the GPU clock)
    e.set_profiling(not args.no_profile)
    t0 = time.perf_counter()
    e.step(DT, args.steps, th, ba)  # returns after the device finished
    t1 = time.perf_counter()
    st = e.stats()
    value = n * args.steps / (t1 - t0)
    ms_per_step = 1e3 * (t1 - t0) / args.steps
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step, {value:.4e} atom-steps/s; stats {st}")
    step_ms, step_n = e.kernel_time("step")
    fin_ms, fin_n = e.kernel_time("finalize")
    e.set_profiling(False)
    avg_step_s = step_ms / max(step_n, 1) * 1e-3
    # mean list length over the timed launches (entries read / atoms / launches)
    n_l = st["list_visits"] / (n * max(step_n, 1)) if step_n else st["mean_list_len"]
    bytes_per_launch = n * (124.0 + 28.0 * n_l)
    achieved = bytes_per_launch / avg_step_s / 1e9 if step_n else 0.0
    log(f"[bench] k_step avg {avg_step_s * 1e6:.2f} us over {step_n} launches, "
        f"k_fin avg {fin_ms / max(fin_n, 1) * 1e3:.2f} us over {fin_n}; <n_l> {n_l:.4f}")
    from moldyn_amd.rocprof_summary import read_traffic
    traffic = read_traffic(args.traffic)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": f"k_step<float,{opts['block_size']}>", "avg_us": round(avg_step_s * 1e6, 3),
                "launches_timed": step_n, "mean_list_len_timed": round(n_l, 5),
                "k_fin_avg_us": round(fin_ms / max(fin_n, 1) * 1e3, 3),
                "algorithmic_bytes_per_launch": round(bytes_per_launch)}

    # FP32 force error on the final state, against the FP64 path (bitwise the
    # reference update_force, see tests/test_gpu_parity.py), both evaluated at
    # the same positions (a step's stored forces precede its barostat scaling)
    d = e.download(("x", "v", "box"))
    e.upload(d["x"], d["v"], d["box"])
    e.update_force()
    f32 = e.download(("force",))["force"]
    e.close()
    ref = M.Engine([n], [MASS], d["box"], force_mode="fp64_list")
    ref.upload(d["x"], d["v"], d["box"])
    ref.update_force()
    f64 = ref.download(("force",))["force"]
    ref.close()
    err = f32 - f64
    rms_f = float(np.sqrt(np.mean(f64 ** 2)))
    force_rmse = float(np.sqrt(np.mean(err ** 2)))
    force_err = {"rmse": force_rmse, "rms_force": rms_f,
                 "rmse_rel": force_rmse / rms_f if rms_f > 0 else 0.0,
                 "max_abs": float(np.abs(err).max()),
                 "atoms_with_force": int(np.count_nonzero(np.abs(f64).sum(1)))}
    log(f"[bench] force error vs fp64: {force_err}")

    plain = None
    if not args.no_plain:
        p = M.Engine([n], [MASS], box, force_mode="fp64_allpairs")
        p.upload(pos, vel, box)
        p.update_force()
        th2, ba2 = controls(M)
        p.step(DT, 1, th2, ba2)
        t0 = time.perf_counter()
        p.step(DT, args.plain_steps, th2, ba2)
        t1 = time.perf_counter()
        p.close()
        pv = n * args.plain_steps / (t1 - t0)
        plain = {"value": pv, "unit": "atom-steps/s", "steps": args.plain_steps,
                 "ms_per_step": 1e3 * (t1 - t0) / args.plain_steps,
                 "speedup": value / pv}
        log(f"[bench] fp64 plain all-pairs: {pv:.4e} atom-steps/s ({plain['ms_per_step']:.1f} ms/step)")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = max(1, min(os.cpu_count() or 1, 16))
        s = O.State([n], [MASS], box, pos=pos, vel=vel)
        tab = O.potential_table(1)
        rows = 16 * threads
        t0 = time.perf_counter()
        O.update_force_rows(s, 0, rows, tab, threads)
        dt_probe = time.perf_counter() - t0
        rows = int(min(n, max(rows, rows * args.cpu_seconds / max(dt_probe, 1e-6))))
        t0 = time.perf_counter()
        O.update_force_rows(s, 0, rows, tab, threads)
        dt_cpu = time.perf_counter() - t0
        cpu_value = rows / dt_cpu  # one step needs all n rows: n / (dt * n / rows)
        cpu = {"value": cpu_value, "unit": "atom-steps/s", "cores": threads, "kind": "port",
               "sample": f"reference update_force (potential.rs:158-216) rows 0..{rows} of the "
                         f"{n}-atom initial state, each row against all {n} atoms, FP64, "
                         f"OpenMP {threads} threads, {dt_cpu:.1f} s; O(N) integrator passes omitted"}
        log(f"[bench] cpu baseline: {cpu_value:.4e} atom-steps/s on {threads} threads")

    out = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair forces + f64 state/reductions", "data": "synthetic",
        "config": {"workload": f"argon SC {args.n_side}^3 ({n} atoms) LJ NPT, Berendsen T 300 K "
                               f"tau 10, Berendsen P 1.01325 (1, 5), dt {DT}, Verlet list",
                   "n_atoms": n, "force_mode": "fp32_list",
                   "skin_nm": opts["skin"], "block_size": opts["block_size"],
                   "cells_per_atom": opts["cells_per_atom"], "autotuned": not args.no_tune,
                   "rebuilds": st["rebuilds"], "mean_list_len": n_l},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_plain": plain,
        "force_error_vs_fp64": force_err,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
