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


def nve_64k(M, steps=10000):
    """64,000-atom argon, SC 40^3 at a = 3.338339 with uniform jitter of
    +-0.02 a (seed 2), MB 120 K (seed 2), NVE, skin 0.3 sigma, FP32 forces:
    energy drift per ns and the force RMSE against the FP64 path."""
    n_side = 40
    n = n_side ** 3
    rng = np.random.default_rng(2)
    pos = M.lattice("u", [n_side] * 3, A_LAT) + rng.uniform(-0.02, 0.02, (n, 3)) * A_LAT
    box = np.full(3, n_side * A_LAT)
    pos = np.mod(pos, box)
    vel = M.mb_velocities(n, 120.0, MASS, 2)
    e = M.Engine([n], [MASS], box, force_mode="fp32_list", skin=0.3 * 0.3418)
    e.upload(pos, vel, box)
    e.update_force()
    m0 = e.macro()
    t0 = time.perf_counter()
    e.step(DT, steps)
    wall = time.perf_counter() - t0
    m1 = e.macro()
    d = e.download(("x", "v", "force", "box"))
    e.close()
    f32 = gpu_forces(M, d["x"], d["box"], "fp32_list")
    f64 = gpu_forces(M, d["x"], d["box"], "fp64_list")
    e0 = m0["kinetic"] + m0["potential"]
    e1 = m1["kinetic"] + m1["potential"]
    ns = steps * DT * 1e-3
    res = {"workload": "argon SC 40^3 (64000 atoms), +-0.02a jitter, 120 K, NVE, skin 0.3 sigma, FP32 forces",
           "steps": steps, "atom_steps_per_s": n * steps / wall,
           "energy_drift_per_ns": (e1 - e0) / ns, "energy_drift_rel_per_ns": (e1 - e0) / abs(e0) / ns,
           "total_energy_start": e0,
           "force_error_final_vs_fp64_list": force_error(f32, f64)}
    return res, {"x": d["x"], "v": d["v"], "box": d["box"], "f32": f32}


def replicas_256(M, steps=2000, seq_sample=48):
    """256 argon replicas in one engine: N_r = round(exp(U[ln 500, ln 32768]))
    (seed 4) on ceil(N^(1/3))^3 SC grids filled in index order, a_r in
    U[3.0, 4.2], Berendsen T0 in U[80, 400] K and P0 in U[0.5, 5.0]; the
    aggregate rate against the same replicas run one engine at a time."""
    rng = np.random.default_rng(4)
    R = 256
    counts = np.rint(np.exp(rng.uniform(np.log(500), np.log(32768), R))).astype(np.int64)
    a = rng.uniform(3.0, 4.2, R)
    t0 = rng.uniform(80.0, 400.0, R)
    p0 = rng.uniform(0.5, 5.0, R)
    pos, vel, boxes = [], [], []
    for r in range(R):
        side = int(np.ceil(counts[r] ** (1.0 / 3.0) - 1e-9))
        x = M.lattice("u", [side] * 3, a[r])[:counts[r]]
        pos.append(x)
        vel.append(M.mb_velocities(int(counts[r]), t0[r], MASS, 1000 + r))
        boxes.append([side * a[r]] * 3)
    boxes = np.array(boxes)
    L = M.engine.L

    def controls():
        th = (L.MdThermostat * R)(*[M.thermostat("berendsen", 10.0, float(t0[r])) for r in range(R)])
        ba = (L.MdBarostat * R)(*[M.barostat("berendsen", 1.0, 5.0, float(p0[r])) for r in range(R)])
        return th, ba

    # skin, block size, tiles per block and cell budget for this batch by the
    # autotuner, on the state half-way through the run (the barostats
    # compress most replicas, which lengthens the lists; dilute gases favour
    # a wide skin, denser ones a thinner)
    from moldyn_amd.autotune import autotune
    xs, vs = np.concatenate(pos), np.concatenate(vel)
    e0 = M.Engine(counts[:, None], [MASS], boxes, force_mode="fp32_list")
    e0.upload(xs, vs, boxes)
    e0.update_force()
    th0, ba0 = controls()
    e0.step(DT, 20 + steps // 2, th0, ba0)
    mid = e0.download(("x", "v", "box"))
    e0.close()
    # 1000-step trials: about ten replicas rebuild per launch, so shorter
    # windows rank the skins by which replicas happened to rebuild in them
    best, _ = autotune(counts[:, None], [MASS], mid["box"], mid["x"], mid["v"], DT, controls, steps=1000,
                       warmup=20, blocks=(128, 256, 512), tiles=(0, 1, 2, 4), cells=(1.0, 2.0, 4.0), log=log)
    e = M.Engine(counts[:, None], [MASS], boxes, force_mode="fp32_list", **best)
    e.upload(xs, vs, boxes)
    e.update_force()
    th, ba = controls()
    e.step(DT, 20, th, ba)
    tb = time.perf_counter()
    e.step(DT, steps, th, ba)
    wall_b = time.perf_counter() - tb
    e.close()
    total = int(counts.sum())
    # one engine per replica, a sample of them, scaled to all atoms
    idx = rng.choice(R, seq_sample, replace=False)
    wall_s, atoms_s = 0.0, 0
    for r in idx:
        one = M.Engine([int(counts[r])], [MASS], boxes[r], force_mode="fp32_list", **best)
        one.upload(pos[r], vel[r], boxes[r])
        one.update_force()
        th1 = M.thermostat("berendsen", 10.0, float(t0[r]))
        ba1 = M.barostat("berendsen", 1.0, 5.0, float(p0[r]))
        one.step(DT, 20, th1, ba1)
        ts = time.perf_counter()
        one.step(DT, steps, th1, ba1)
        wall_s += time.perf_counter() - ts
        atoms_s += int(counts[r])
        one.close()
    seq_rate = atoms_s * steps / wall_s
    return {"workload": "256 argon replicas, N log-uniform in [500, 32768] (seed 4), a in [3.0, 4.2], "
                        "Berendsen T0 in [80, 400] K, P0 in [0.5, 5.0], FP32 forces",
            "autotuned": best,
            "atoms": total, "steps": steps, "batched_atom_steps_per_s": total * steps / wall_b,
            "batched_ms_per_step": 1e3 * wall_b / steps,
            "sequential_atom_steps_per_s": seq_rate,
            "sequential_sample": f"{seq_sample} replicas ({atoms_s} atoms) one engine each, timed steps only",
            "speedup": total * steps / wall_b / seq_rate}


def frame_io(M, pos, vel, box, frames=6, steps_between=20):
    """State files of the 1M-atom state (data/<n>.csv + bb.csv in the
    reference's exact text): GPU text kernels per frame, the writer's
    end-to-end rate with saves back to back, and the cost of saving every
    `steps_between` steps next to stepping alone."""
    import shutil
    import tempfile
    from moldyn_amd.frames import FrameWriter
    n = len(pos)
    tmp = tempfile.mkdtemp(prefix="md_frames_")
    try:
        e = M.Engine([n], [MASS], box, force_mode="fp32_list")
        e.upload(pos, vel, box)
        e.update_force()
        th, ba = controls(M)
        e.step(DT, 2, th, ba)
        with FrameWriter(e, tmp) as w:
            w.save(0)
            w.flush()
            t0 = time.perf_counter()
            for k in range(1, frames + 1):
                w.save(k)
            w.flush()
            wall_b2b = time.perf_counter() - t0
            st = w.stats()
        size = os.path.getsize(os.path.join(tmp, "data", "1.csv"))
        # stepping alone, then stepping with a save every `steps_between` steps
        ts = time.perf_counter()
        for _ in range(frames):
            e.step(DT, steps_between, th, ba)
        plain = time.perf_counter() - ts
        with FrameWriter(e, tmp) as w:
            ts = time.perf_counter()
            for k in range(frames):
                w.save(100 + k)
                e.step(DT, steps_between, th, ba)
            w.flush()
            with_io = time.perf_counter() - ts
        e.close()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    gpu_ms = st["gpu_ms"] / st["frames"]
    return {"workload": f"{n}-atom state as data/<n>.csv + bb.csv (reference text), written to local disk",
            "frame_bytes": size, "frames": frames,
            "gpu_text_ms_per_frame": gpu_ms, "gpu_text_GBps": size / gpu_ms / 1e6,
            "copy_ms_per_frame": st["copy_ms"] / st["frames"], "write_ms_per_frame": st["write_ms"] / st["frames"],
            "back_to_back_frames_per_s": frames / wall_b2b, "back_to_back_MBps": frames * size / wall_b2b / 1e6,
            f"steps_{steps_between}_plain_s": plain, f"steps_{steps_between}_with_save_s": with_io}


def liquid_262k_state(M):
    """64^3 SC argon at a = 0.47 nm (262,144 atoms) equilibrated 500 steps at
    150 K by the FP32 LJ path (Berendsen T, tau 0.1 ps): (n, a, state, mean
    pressure over the last 100 steps)."""
    n_side, a = 64, 0.47
    n = n_side ** 3
    box = np.full(3, n_side * a)
    rng = np.random.default_rng(6)
    pos = np.mod(M.lattice("u", [n_side] * 3, a) + rng.uniform(-0.02, 0.02, (n, 3)) * a, box)
    vel = M.mb_velocities(n, 150.0, MASS, 6)
    probe = M.Engine([n], [MASS], box, force_mode="fp32_list")
    probe.upload(pos, vel, box)
    probe.update_force()
    th_eq = M.thermostat("berendsen", tau=0.1, target_t=150.0)
    probe.step(DT, 400, th_eq)
    p_hist = []
    for _ in range(5):
        probe.step(DT, 20, th_eq)
        p_hist.append(probe.macro()["pressure"])
    st = probe.download(("x", "v", "box"))
    probe.close()
    return n, a, st, float(np.mean(p_hist))


def dense_liquid(M, steps):
    """The FP32 LJ path alone on the equilibrated 262,144-atom liquid under
    Berendsen T and P (the surrogate workload's LJ leg), with per-launch
    device timing: the dense-list force case for hardware counters."""
    n, a, st, p_state = liquid_262k_state(M)
    e = M.Engine([n], [MASS], st["box"], force_mode="fp32_list")
    e.upload(st["x"], st["v"], st["box"])
    e.update_force()
    th = M.thermostat("berendsen", tau=1.0, target_t=150.0)
    ba = M.barostat("berendsen", beta=1.0, tau=50.0, target_p=p_state)
    e.step(DT, 20, th, ba)
    e.set_profiling(True)
    ts = time.perf_counter()
    e.step(DT, steps, th, ba)
    mac = e.macro()
    wall = time.perf_counter() - ts
    step_ms, step_n = e.kernel_time("step")
    fin_ms, fin_n = e.kernel_time("finalize")
    span_ms, span_n = e.kernel_time("step", span=True)
    stt = e.stats()
    e.close()
    return {"workload": f"64^3 SC argon ({n} atoms) at a = {a} nm, FP32 LJ list, Berendsen T and P, dt {DT}, "
                        f"{steps} steps",
            "atom_steps_per_s": n * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "k_step_avg_us": 1e3 * step_ms / max(step_n, 1), "k_fin_avg_us": 1e3 * fin_ms / max(fin_n, 1),
            "k_step_span_avg_us": 1e3 * span_ms / max(span_n, 1),
            "list_entries_per_launch": stt["list_visits"] / max(span_n, 1),
            "mean_list_len": stt["list_visits"] / max(span_n, 1) / n, "rebuilds": stt["rebuilds"],
            "temperature_final": mac["temperature"]}


def mlp_surrogate(M, n_frames=2000, epochs=40, steps=2000, log=log, prof_steps=200):
    """Surrogate: 4,096-atom LJ argon frames from thermostatted SC runs at
    lattice edges 0.38-0.7 nm and 80-400 K, labelled by the FP64 LJ path,
    trained on the GPU; inference on a 64^3 (262,144-atom) SC grid at
    a = 0.47 nm equilibrated at 150 K, then Berendsen T and P at the
    equilibrated pressure, against the FP32 LJ path on the same run."""
    from moldyn_amd import mlp
    t0 = time.perf_counter()
    fr = mlp.make_frames(n_side=16, n_frames=n_frames, runs=20, equil=400, every=20, seed=5)
    t_gen = time.perf_counter() - t0
    held = np.arange(n_frames - 50, n_frames, 5)  # every 5th of the last run's frames, held out
    train_idx = np.setdiff1d(np.arange(n_frames), np.arange(n_frames - 50, n_frames))
    t0 = time.perf_counter()
    sur = mlp.train(fr.subset(train_idx), epochs=epochs, batch_frames=16, lr=2e-3, lr_final=1e-4, log=log)
    t_train = time.perf_counter() - t0
    # held-out force error of the deployed (BF16 MFMA) surrogate
    rel = []
    for k in held:
        x, box = fr.pos[k], fr.box[k]
        e = M.Engine([len(x)], [MASS], box, force_mode="mlp")
        sur.apply(e, judge_every=0, range_check=False)
        e.upload(x, np.zeros_like(x), box)
        e.update_force()
        f = e.download(("force",))["force"]
        e.close()
        rel.append(float(np.sqrt(np.mean((f - fr.f_ref[k]) ** 2)) / np.sqrt(np.mean(fr.f_ref[k] ** 2))))
    n, a, st, p_state = liquid_262k_state(M)
    res = {}
    for mode in ("mlp", "fp32_list"):
        e = M.Engine([n], [MASS], st["box"], force_mode=mode)
        if mode == "mlp":
            sur.apply(e, judge_every=100, range_check=True)
        e.upload(st["x"], st["v"], st["box"])
        e.update_force()
        th = M.thermostat("berendsen", tau=1.0, target_t=150.0)
        ba = M.barostat("berendsen", beta=1.0, tau=50.0, target_p=p_state)
        e.step(DT, 20, th, ba)
        ts = time.perf_counter()
        e.step(DT, steps, th, ba)
        wall = time.perf_counter() - ts
        st_run = e.stats()
        mac = e.macro()
        res[mode] = {"atom_steps_per_s": n * steps / wall, "ms_per_step": 1e3 * wall / steps,
                     "rebuilds": st_run["rebuilds"], "launches": st_run["launches"],
                     "mean_list_len": st_run["mean_list_len"],
                     "temperature_final": mac["temperature"], "pressure_final": mac["pressure"],
                     "box_final": float(e.download(("box",))["box"][0])}
        if mode == "mlp":
            res[mode].update(e.mlp_stats())
        # per-kernel device time over a further profiled stretch
        e.set_profiling(True)
        e.step(DT, prof_steps, th, ba)
        kern = {}
        for name, key in (("k_mlp_desc" if mode == "mlp" else None, "drift"),
                          ("k_mlp_feat" if mode == "mlp" else None, "mlp_feat"),
                          ("k_mlp_net" if mode == "mlp" else None, "mlp_net"),
                          ("k_mlp_force" if mode == "mlp" else "k_step", "step"), ("k_fin", "finalize")):
            if name is None:
                continue
            ms, cnt = e.kernel_time(key)
            sms, scnt = e.kernel_time(key, span=True)
            kern[name] = {"stream_us": 1e3 * ms / max(cnt, 1), "span_us": 1e3 * sms / max(scnt, 1),
                          "launches": cnt}
        e.set_profiling(False)
        res[mode]["kernels"] = kern
        if mode == "mlp":
            # 80 MFMA 16x16x32 (2*16*16*32 flops each) per 16 atoms: the two
            # forward layers (8 + 32), the backward through W2 (32) and W1 (8)
            flops = n * 80 * 2 * 16 * 16 * 32 / 16
            t_net = kern["k_mlp_net"]["span_us"] * 1e-6
            res[mode]["k_mlp_net_mfma"] = {
                "flops_per_launch": flops, "span_us": kern["k_mlp_net"]["span_us"],
                "achieved_tflops": flops / t_net / 1e12 if t_net > 0 else None,
                "peak_tflops_bf16_dense": 2500.0,
                "frac": flops / t_net / 2.5e15 if t_net > 0 else None,
                "basis": "device timestamps (s_memrealtime) of every block of every launch, first start to "
                         "last end"}
        e.close()
    return {"workload": f"64^3 SC argon ({n} atoms) at a = {a} nm equilibrated 500 steps at 150 K, then Berendsen T "
                        f"(150 K, tau 1) and P (beta 1, tau 50, P0 = equilibrated pressure {p_state:.4g}), "
                        f"dt {DT}, {steps} steps; box at start {float(st['box'][0]):.4f}",
            "training": {"frames": int(len(train_idx)), "atoms_per_frame": 4096, "epochs": epochs,
                         "frame_generation_s": t_gen, "training_s": t_train,
                         "final_loss": sur.history[-1], "first_loss": sur.history[0]},
            "heldout_force_rmse_rel": float(np.mean(rel)),
            "mlp": res["mlp"], "lj_fp32": res["fp32_list"],
            "speed_vs_lj": res["mlp"]["atom_steps_per_s"] / res["fp32_list"]["atom_steps_per_s"]}


def kernel_roofline(e, n, steps, block_size, traffic):
    """Roofline of k_step from a profiled pass: stream time per launch on the
    device clock, algorithmic bytes A = n (96 + 28 <n_l>) per launch."""
    step_ms, step_n = e.kernel_time("step")
    fin_ms, fin_n = e.kernel_time("finalize")
    span_ms, span_n = e.kernel_time("step", span=True)
    st = e.stats()
    avg_s = step_ms / max(step_n, 1) * 1e-3
    # mean list length over the launches (entries read / atoms / launches)
    n_l = st["list_visits"] / (n * max(span_n, 1)) if span_n else st["mean_list_len"]
    # step records in and out (48 B each), 4 B index + 24 B position per list entry
    bytes_per_launch = n * (96.0 + 28.0 * n_l)
    achieved = bytes_per_launch / avg_s / 1e9 if step_n else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": f"k_step<float,{block_size}>", "avg_us": round(avg_s * 1e6, 3),
            "avg_us_basis": "stream time: previous kernel end to kernel end on the GPU clock, back-to-back "
                            "launches (the interval rocprofv3 reports)",
            "span_avg_us": round(span_ms / max(span_n, 1) * 1e3, 3), "launches_timed": step_n,
            "steps": steps, "mean_list_len": round(n_l, 5),
            "k_fin_avg_us": round(fin_ms / max(fin_n, 1) * 1e3, 3),
            "algorithmic_bytes_per_launch": round(bytes_per_launch)}


def counter_evidence():
    """Hardware-counter figures from the committed rocprofv3 summaries
    (profiles/): HBM GB/s of the neighbour build and of integration at
    1M atoms, and the VALU FLOP rates of the force loop on the dense liquid
    (the FLOPS counters count per wave instruction: x 64 lanes)."""
    from rocprof_summary import read_counters, read_rate
    out = {}
    prof = os.path.join(ROOT, "profiles")
    for key, name in (("integration_k_step_1M", "k_step_1M_npt_traffic.txt"),
                      ("neighbour_build_k_fin_1M", "k_fin_build_1M_npt_traffic.txt")):
        r = read_rate(os.path.join(prof, name))
        if r and r[0] and r[1]:
            gbs = r[0] / (r[1] * 1e-6) / 1e9
            out[key] = {"bytes_per_launch": r[0], "median_us": r[1], "hbm_GBps": round(gbs, 1),
                        "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4), "source": f"profiles/{name}"}
    c = read_counters(os.path.join(prof, "k_step_262k_liquid_counters.txt"))
    if c and c[1]:
        v, dur = c
        f32 = v.get("SQ_INSTS_VALU_FLOPS_FP32", 0.0) * 64 / (dur * 1e-6)
        f64 = v.get("SQ_INSTS_VALU_FLOPS_FP64", 0.0) * 64 / (dur * 1e-6)
        out["forces_k_step_262k_liquid"] = {
            "median_us": dur, "fp32_TFLOPs": round(f32 / 1e12, 3), "fp64_TFLOPs": round(f64 / 1e12, 3),
            "fp32_frac_of_157.3": round(f32 / 157.3e12, 4), "fp64_frac_of_78.6": round(f64 / 78.6e12, 4),
            "valu_wave_insts": v.get("SQ_INSTS_VALU"), "source": "profiles/k_step_262k_liquid_counters.txt"}
    c = read_counters(os.path.join(prof, "k_step_1M_npt_counters.txt"))
    if c and c[0].get("SQ_WAVE_CYCLES"):
        v = c[0]
        wc = v["SQ_WAVE_CYCLES"]
        tiles = 1_000_000 / 64  # wave-tiles of 64 atoms per launch
        out["integration_k_step_1M_sq"] = {
            "wait_any_frac": round(v.get("SQ_WAIT_ANY", 0.0) / wc, 4),
            "wait_inst_frac": round(v.get("SQ_WAIT_INST_ANY", 0.0) / wc, 4),
            "valu_active_frac": round(v.get("SQ_ACTIVE_INST_VALU", 0.0) / wc, 4),
            "valu_insts_per_wave_tile": round(v.get("SQ_INSTS_VALU", 0.0) / tiles, 1),
            "salu_insts_per_wave_tile": round(v.get("SQ_INSTS_SALU", 0.0) / tiles, 1),
            "vmem_rd_per_wave_tile": round(v.get("SQ_INSTS_VMEM_RD", 0.0) / tiles, 2),
            "vmem_wr_per_wave_tile": round(v.get("SQ_INSTS_VMEM_WR", 0.0) / tiles, 2),
            "source": "profiles/k_step_1M_npt_counters.txt"}
    return out or None


def force_error(f, ref):
    """Force error of f against ref (n x 3); relative figures are None when
    no atom of ref has a force (the comparison would be vacuous)."""
    err = f - ref
    with_force = int(np.count_nonzero(np.abs(ref).sum(1)))
    rms_f = float(np.sqrt(np.mean(ref ** 2)))
    rmse = float(np.sqrt(np.mean(err ** 2)))
    out = {"rmse": rmse, "rms_force": rms_f, "max_abs": float(np.abs(err).max()),
           "atoms_with_force": with_force, "atoms": len(ref)}
    if with_force == 0:
        out.update({"rmse_rel": None, "max_rel": None, "degenerate": True})
    else:
        out.update({"rmse_rel": rmse / rms_f, "max_rel": float(np.abs(err).max()) / rms_f, "degenerate": False})
    return out


def gpu_forces(M, x, box, mode):
    n = len(x)
    e = M.Engine([n], [MASS], box, force_mode=mode)
    e.upload(x, np.zeros_like(x), box)
    e.update_force()
    f = e.download(("force",))["force"]
    e.close()
    return f


def cpu_leg(M, args, n, pos, vel, box, final, nve_final):
    """The CPU side, on the host cores of the GPU's share: the FP64 oracle
    (a restatement of the reference solver) timed as the baseline, its
    cell-grid variant beside it, and the oracle as the checker of the GPU
    results (forces on states with pairs, config-1 mean T and P, 100-step
    trajectory deviation)."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    host = {"host_cpus_visible": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "threads": threads}
    tab = O.potential_table(1)
    # the reference algorithm: update_force rows (potential.rs:158-216), a
    # bounded sample of the O(N^2) step, scaled to atom-steps/s
    s = O.State([n], [MASS], box, pos=pos, vel=vel)
    rows = 16 * threads
    t0 = time.perf_counter()
    O.update_force_rows(s, 0, rows, tab, threads)
    dt_probe = time.perf_counter() - t0
    rows = int(min(n, max(rows, rows * args.cpu_seconds / max(dt_probe, 1e-6))))
    t0 = time.perf_counter()
    O.update_force_rows(s, 0, rows, tab, threads)
    dt_cpu = time.perf_counter() - t0
    cpu_value = rows / dt_cpu  # one step needs all n rows: n / (dt * n / rows)
    log(f"[bench] cpu baseline: {cpu_value:.4e} atom-steps/s on {threads} threads")
    # the same FP64 arithmetic through a cell grid (bitwise the same forces),
    # whole integrator steps under the thermostat and barostat
    cs = O.State([n], [MASS], box, pos=pos, vel=vel)
    O.update_force_cells(cs, tab, threads)
    oth, oba = O.Thermostat.berendsen(10.0), O.Barostat.berendsen(1.0, 5.0)
    O.verlet_step(cs, DT, tab, oba, 1.01325, oth, 300.0, threads, cells=True)
    t0 = time.perf_counter()
    for _ in range(args.cpu_cell_steps):
        O.verlet_step(cs, DT, tab, oba, 1.01325, oth, 300.0, threads, cells=True)
    dt_cells = time.perf_counter() - t0
    cell_value = n * args.cpu_cell_steps / dt_cells
    log(f"[bench] cpu cell-list variant: {cell_value:.4e} atom-steps/s")
    cpu = {"value": cpu_value, "unit": "atom-steps/s", "cores": threads, "kind": "port",
           "sample": f"reference update_force (potential.rs:158-216) rows 0..{rows} of the {n}-atom initial "
                     f"state, each row against all {n} atoms, FP64, OpenMP {threads} threads, {dt_cpu:.1f} s; "
                     f"the O(N) integrator passes (1e-6 of this work at this size) omitted",
           **host,
           "cell_list_variant": {"value": cell_value, "unit": "atom-steps/s", "steps": args.cpu_cell_steps,
                                 "sample": f"{args.cpu_cell_steps} whole Berendsen NPT steps of the {n}-atom "
                                           f"state (integrator.rs:14-64) with the forces through a cell grid, "
                                           f"FP64, OpenMP {threads} threads, {dt_cells:.2f} s"}}

    acc = {}
    # forces of the bench's own final state, FP32 GPU against the oracle
    fs = O.State([n], [MASS], final["box"], pos=final["x"], vel=final["v"])
    O.update_force_cells(fs, tab, threads)
    acc["force_1M_final_state"] = {"state": f"final state of the {args.full_steps}-step run, box "
                                            f"{final['box'][0]:.4f} nm",
                                   **force_error(final["f32"], fs.force)}
    acc["force_1M_final_state"]["fp64_list_bitwise"] = bool(np.array_equal(final["f64"], fs.force))
    # dense fixtures: FCC liquid argon (rho sigma^3 ~ 0.8) and a gas at ten
    # times the workload's density
    rng = np.random.default_rng(5)
    liq = O.lattice_fcc([8, 8, 8], 0.5256)
    lbox = np.full(3, 8 * 0.5256)
    liq = np.mod(liq + rng.uniform(-0.02, 0.02, liq.shape), lbox)
    ls = O.State([len(liq)], [MASS], lbox, pos=liq)
    O.update_force(ls, tab, threads)
    acc["force_liquid_fcc"] = {"state": "FCC argon 8^3 cells (2048 atoms), a = 0.5256 nm, +-0.02 nm jitter",
                               **force_error(gpu_forces(M, liq, lbox, "fp32_list"), ls.force)}
    a10 = A_LAT / 10.0 ** (1.0 / 3.0)
    gas = O.lattice_u([30, 30, 30], a10)
    gbox = np.full(3, 30 * a10)
    gas = np.mod(gas + rng.uniform(-0.25, 0.25, gas.shape) * a10, gbox)
    gs = O.State([len(gas)], [MASS], gbox, pos=gas)
    O.update_force_cells(gs, tab, threads)
    acc["force_gas_10x"] = {"state": f"27000 argon atoms at ten times the workload's density (SC a = {a10:.4f} nm, "
                                     f"+-0.25 a jitter)",
                            **force_error(gpu_forces(M, gas, gbox, "fp32_list"), gs.force)}
    if nve_final is not None:
        ns = O.State([len(nve_final["x"])], [MASS], nve_final["box"], pos=nve_final["x"], vel=nve_final["v"])
        O.update_force_cells(ns, tab, threads)
        acc["force_64k_nve_final_state"] = {"state": "final state of the 10,000-step 64k NVE run",
                                            **force_error(nve_final["f32"], ns.force)}
    # config 1: 1,000 atoms, Berendsen T and P, 1,000 steps; the FP32 path's
    # temperature and pressure every step against the oracle's
    c1 = O.lattice_u([10, 10, 10], A_LAT)
    c1v = M.mb_velocities(1000, 273.15, MASS, 1)
    c1box = np.full(3, 10 * A_LAT)
    e = M.Engine([1000], [MASS], c1box, force_mode="fp32_list")
    e.upload(c1, c1v, c1box)
    e.update_force()
    ref = O.State([1000], [MASS], c1box, pos=c1, vel=c1v)
    O.apply_boundary_conditions(ref)
    O.update_force(ref, tab, threads)
    th, ba = controls(M)
    oth, oba = O.Thermostat.berendsen(10.0), O.Barostat.berendsen(1.0, 5.0)
    tg, pg, to, po, dx100 = [], [], [], [], None
    for k in range(1000):
        e.step(DT, 1, th, ba)
        O.verlet_step(ref, DT, tab, oba, 1.01325, oth, 300.0, threads)
        m = e.macro()
        mv = O.center_of_mass_velocity(ref, 0)
        tg.append(m["temperature"])
        pg.append(m["pressure"])
        to.append(O.temperature(O.thermal_energy(ref, 0, mv), 1000))
        po.append(O.pressure(ref, 0, mv))
        if k == 99:
            d = e.download(("x",))["x"] - ref.pos
            d -= ref.box * np.round(d / ref.box)
            dx100 = float(np.abs(d).max())
    fin = e.download(("x", "box"))
    e.close()
    d = fin["x"] - ref.pos
    d -= ref.box * np.round(d / ref.box)
    tg, pg, to, po = map(np.array, (tg, pg, to, po))
    acc["config1_npt_1000_steps"] = {
        "path": "fp32_list", "mean_T": float(tg.mean()), "mean_T_oracle": float(to.mean()),
        "mean_P": float(pg.mean()), "mean_P_oracle": float(po.mean()),
        "max_abs_dT": float(np.abs(tg - to).max()), "max_rel_dP": float(np.abs(pg - po).max() / np.abs(po).max()),
        "max_dx_100_steps": dx100, "max_dx_1000_steps": float(np.abs(d).max()),
        "box_rel_diff": float(np.abs(fin["box"] - ref.box).max() / ref.box[0])}
    # 100-step trajectory deviation where forces act: the liquid, NVE
    lv = M.mb_velocities(len(liq), 120.0, MASS, 5)
    e = M.Engine([len(liq)], [MASS], lbox, force_mode="fp32_list")
    e.upload(liq, lv, lbox)
    e.update_force()
    e.step(DT, 100)
    got = e.download(("x", "v"))
    e.close()
    ref = O.State([len(liq)], [MASS], lbox, pos=liq, vel=lv)
    O.update_force(ref, tab, threads)
    for _ in range(100):
        O.verlet_step(ref, DT, tab, nthreads=threads)
    d = got["x"] - ref.pos
    d -= lbox * np.round(d / lbox)
    acc["trajectory_100_steps_liquid"] = {
        "state": "FCC argon liquid (2048 atoms), 120 K, NVE, FP32 path against the FP64 oracle",
        "max_dx_nm": float(np.abs(d).max()), "rms_dx_nm": float(np.sqrt(np.mean(d ** 2))),
        "max_dv_rel": float(np.abs(got["v"] - ref.vel).max() / np.sqrt(np.mean(ref.vel ** 2)))}
    log(f"[bench] accuracy vs oracle: {acc}")
    return cpu, acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--full-steps", type=int, default=5000,
                    help="length of the workload's full run, reported beside the timed steps")
    ap.add_argument("--n-side", type=int, default=N_SIDE)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-cell-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines and the oracle checks")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--skin", type=float, default=0.5, help="nm (the autotuner's choice for this workload)")
    ap.add_argument("--block-size", type=int, default=256)
    ap.add_argument("--cells-per-atom", type=float, default=4.0)
    ap.add_argument("--tiles-per-block", type=int, default=0)
    ap.add_argument("--graph-chunk", type=int, default=32, help="launches per captured graph (1: none)")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-launch device timestamps (no roofline)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tune-steps", type=int, default=1000)
    ap.add_argument("--tune-repeats", type=int, default=2)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 64k NVE, 256-replica and surrogate measurements")
    ap.add_argument("--mlp-frames", type=int, default=2000)
    ap.add_argument("--mlp-epochs", type=int, default=40)
    ap.add_argument("--only", choices=("mlp", "liquid"), default=None,
                    help="run one secondary workload alone and print its JSON (for profiling): 'mlp', the "
                         "learned surrogate's training and 262,144-atom liquid run beside the FP32 LJ path; "
                         "'liquid', that FP32 LJ liquid run alone")
    ap.add_argument("--mlp-steps", type=int, default=2000)
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "k_step_1M_npt_traffic.txt"),
                    help="rocprof_summary traffic file holding bytes_per_launch of the step kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(this engine runs one system or replica batch per GPU)", file=sys.stderr)
        sys.exit(2)

    import moldyn_amd as M
    from rocprof_summary import read_traffic
    traffic = read_traffic(args.traffic)
    if args.only == "mlp":
        res = mlp_surrogate(M, n_frames=args.mlp_frames, epochs=args.mlp_epochs, steps=args.mlp_steps)
        print(json.dumps({"workload": "mlp_surrogate", **res}, default=float), flush=True)
        return
    if args.only == "liquid":
        res = dense_liquid(M, args.mlp_steps)
        print(json.dumps({"workload": "dense_liquid", **res}, default=float), flush=True)
        return

    n, pos, vel, box = initial_state(M, args.n_side)
    log(f"[bench] state ready: {n} atoms, box {box[0]:.4f}")
    opts = {"skin": args.skin,
            "block_size": args.block_size, "cells_per_atom": args.cells_per_atom,
            "tiles_per_block": args.tiles_per_block}
    full_steps = max(args.full_steps, args.steps)
    if not args.no_tune:
        from moldyn_amd.autotune import autotune
        # tune on the state half-way through the full run: the barostat
        # compresses the gas several-fold over it, which lengthens the lists
        # and shortens the rebuild interval
        pre = args.warmup + full_steps // 2
        e0 = M.Engine([n], [MASS], box, force_mode="fp32_list", skin=0.5, cells_per_atom=4.0)
        e0.upload(pos, vel, box)
        e0.update_force()
        th0, ba0 = controls(M)
        e0.step(DT, pre, th0, ba0)
        mid = e0.download(("x", "v", "box"))
        e0.close()
        log(f"[bench] tuning at step {pre}: box {mid['box'][0]:.4f}")
        # 1000-step windows hold about ten list builds at this state, and
        # the faster of two windows counts: settings within a few per cent
        # of each other (the cell budget, tiles per block) are otherwise
        # ranked by noise
        opts, _ = autotune([n], [MASS], mid["box"], mid["x"], mid["v"], DT, lambda: controls(M),
                           steps=args.tune_steps, warmup=20, log=log, repeats=args.tune_repeats)
        log(f"[bench] autotuned options: {opts}")
    e = M.Engine([n], [MASS], box, force_mode="fp32_list", graph_chunk=args.graph_chunk, **opts)

    def run(steps, profile):
        # warmup then the timed steps, from the initial state
        e.upload(pos, vel, box)
        e.update_force()
        th, ba = controls(M)
        e.step(DT, args.warmup, th, ba)
        e.set_profiling(profile)
        t0 = time.perf_counter()
        e.step(DT, steps, th, ba)  # returns after the device finished
        t1 = time.perf_counter()
        st = e.stats()
        if profile:
            roof = kernel_roofline(e, n, steps, opts["block_size"], traffic)
            e.set_profiling(False)
            return t1 - t0, st, roof
        return t1 - t0, st, None

    # the timed region: exactly --steps steps
    wall, st, _ = run(args.steps, False)
    value = n * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step, {value:.4e} atom-steps/s; stats {st}")
    timed_roof = None
    if not args.no_profile:
        # the same steps again with per-launch device timestamps (the graphs
        # keep running; the results are bitwise those of the timed pass)
        _, _, timed_roof = run(args.steps, True)
        log(f"[bench] timed region k_step {timed_roof['avg_us']} us, <n_l> {timed_roof['mean_list_len']}")
    # the workload's full run (5,000 steps), whatever --steps was
    if full_steps == args.steps:
        wall_f, st_f = wall, st
    else:
        wall_f, st_f, _ = run(full_steps, False)
    final = e.download(("x", "v", "box"))
    full_roof = None
    if not args.no_profile:
        if full_steps == args.steps:
            full_roof = timed_roof
        else:
            _, _, full_roof = run(full_steps, True)
    full_run = {"steps": full_steps, "atom_steps_per_s": n * full_steps / wall_f,
                "ms_per_step": 1e3 * wall_f / full_steps, "rebuilds": st_f["rebuilds"],
                "box_final": float(final["box"][0]), "roofline": full_roof}
    log(f"[bench] full run: {full_run}")
    # the representative roofline is the full run's; the timed region's beside it
    roofline = dict(full_roof) if full_roof else {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS,
                                                  "unit": "GB/s", "frac": None, "traffic": traffic}
    roofline["basis"] = f"k_step over the full {full_steps}-step run (the workload), device clock"
    roofline["timed_region"] = ({k: timed_roof[k] for k in ("achieved", "frac", "avg_us", "mean_list_len",
                                                             "launches_timed", "k_fin_avg_us")}
                                if timed_roof else None)

    # forces of the final state at identical positions: the FP32 path and the
    # FP64 list path (bitwise the reference update_force)
    final["f32"] = gpu_forces(M, final["x"], final["box"], "fp32_list")
    final["f64"] = gpu_forces(M, final["x"], final["box"], "fp64_list")
    e.close()
    force_err = {"state": f"final state of the {full_steps}-step run", **force_error(final["f32"], final["f64"]),
                 "reference": "FP64 list path on the GPU (bitwise the oracle's update_force)"}
    log(f"[bench] force error vs fp64 list: {force_err}")

    comparators = {}
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
        comparators["fp64_plain"] = {"value": pv, "unit": "atom-steps/s", "steps": args.plain_steps,
                                     "ms_per_step": 1e3 * (t1 - t0) / args.plain_steps, "speedup": value / pv,
                                     "what": "FP64 all-pairs restatement of update_force (potential.rs:158-216), "
                                             "no list, same NPT controls"}
        log(f"[bench] fp64 plain all-pairs: {pv:.4e} atom-steps/s ({comparators['fp64_plain']['ms_per_step']:.1f} ms/step)")
    # the same list machinery in FP64 throughout (reference precision)
    q = M.Engine([n], [MASS], box, force_mode="fp64_list", graph_chunk=args.graph_chunk, **opts)
    q.upload(pos, vel, box)
    q.update_force()
    th3, ba3 = controls(M)
    q.step(DT, args.warmup, th3, ba3)
    t0 = time.perf_counter()
    q.step(DT, full_steps, th3, ba3)
    t1 = time.perf_counter()
    q.close()
    qv = n * full_steps / (t1 - t0)
    comparators["fp64_list"] = {"value": qv, "unit": "atom-steps/s", "steps": full_steps,
                                "ms_per_step": 1e3 * (t1 - t0) / full_steps,
                                "speedup_of_fp32": full_run["atom_steps_per_s"] / qv,
                                "what": "the cell-list / Verlet-skin path with FP64 pair forces (bitwise the "
                                        "reference), full run"}
    log(f"[bench] fp64 list: {qv:.4e} atom-steps/s")

    extras = {}
    nve_final = None
    if not args.no_extras:
        log("[bench] 64k NVE")
        extras["nve_64k"], nve_final = nve_64k(M)
        log(f"[bench] nve_64k: {extras['nve_64k']}")
        log("[bench] 256 replicas")
        extras["replicas_256"] = replicas_256(M)
        log(f"[bench] replicas_256: {extras['replicas_256']}")
        log("[bench] frame files")
        extras["frame_io_1M"] = frame_io(M, pos, vel, box)
        log(f"[bench] frame_io_1M: {extras['frame_io_1M']}")
        log("[bench] surrogate")
        extras["mlp_surrogate"] = mlp_surrogate(M, n_frames=args.mlp_frames, epochs=args.mlp_epochs)
        log(f"[bench] mlp_surrogate: {extras['mlp_surrogate']}")

    cpu, acc = (None, None) if args.no_cpu else cpu_leg(M, args, n, pos, vel, box, final, nve_final)

    out = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair forces + f64 state/reductions", "data": "synthetic",
        "config": {"workload": f"argon SC {args.n_side}^3 ({n} atoms) LJ NPT, Berendsen T 300 K "
                               f"tau 10, Berendsen P 1.01325 (1, 5), dt {DT}, Verlet list",
                   "n_atoms": n, "force_mode": "fp32_list",
                   "skin_nm": opts["skin"], "block_size": opts["block_size"],
                   "tiles_per_block": opts["tiles_per_block"],
                   "cells_per_atom": opts["cells_per_atom"], "autotuned": not args.no_tune,
                   "rebuilds": st["rebuilds"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "full_run": full_run,
        "comparators": comparators,
        "force_error_vs_fp64": force_err,
        "accuracy_vs_oracle": acc,
        "counter_evidence": counter_evidence(),
        "other_workloads": extras or None,
    }
    print(json.dumps(out, default=float), flush=True)


if __name__ == "__main__":
    main()
