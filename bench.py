#!/usr/bin/env python3
"""Benchmark: plain FP32 colony step at 1,048,576 particles (3D periodic mixed LJ / soft-core
bulk, lognormal diameters, packing fraction 0.6), one MI355X.

Prints ONE JSON line (rank 0) with particle-steps/s in steady state (the median of at least
five repeats of at least one second each), the roofline of the dominant kernel (its average
launch time over the same repeats, measured with HIP events inside the library on the step
stream) and the CPU baseline (the oracle's reference-structured step on a bounded sample of
the same relaxed state, timed on this host).  The K steps of --steps are timed as well and
reported beside the steady state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cil_cuda_amd import abi, configs, surrogate  # noqa: E402
from cil_cuda_amd.colony import HipMD  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3
N_PARTICLES = 1 << 20

# Algorithmic bytes per particle per launch (external inputs once + outputs once) of the
# pair kernel in 3D FP32: read sorted x,y,z,sigma (16) + tag (4); write force (12) + pressure (4).
PAIR_BYTES_PER_PARTICLE = 36
# Whole FP32 3D step: update+keys 72, scan 4, scatter+binsort 60, pair 36
STEP_BYTES_PER_PARTICLE = 200


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def relaxed_state(n, seed=1):
    return configs.relaxed_bulk_3d(n, seed=seed, precision=abi.PREC_FP32)


def cpu_baseline(seed=1, n_sample=1 << 17):
    """Time the oracle's reference-structured step (COM link list, per-cell-pair element
    loops, -O3 -ffast-math like the reference's -Ofast) on this host, on a bounded sample:
    a 131,072-particle bulk drawn and relaxed exactly like the benchmark workload (same
    density, diameter distribution, law mix).  All host cores (race-free gather) and 1 core."""
    from oracle import oracle as O

    p, box, col = relaxed_state(n_sample, seed)
    q = abi.Params.from_buffer_copy(p)
    q.precision = abi.PREC_FP64
    # 16 threads: the host-CPU share of one GPU on the measurement machine (the same cap as the
    # OMP_NUM_THREADS it exports); nproc and the CPU model are recorded beside it
    threads = max(1, min(16, os.cpu_count() or 1))
    res = {}
    for nth, nsteps in ((threads, 30), (1, 4)):
        orc = O.Oracle(q, box, fast=True, traversal=0, threads=nth)
        c = col.copy()
        orc.forces(c, 0)
        t0 = time.perf_counter()
        orc.step(c, nsteps, 0)
        dt = time.perf_counter() - t0
        orc.close()
        res[nth] = (n_sample * nsteps / dt, nsteps, dt)
        log(f"cpu baseline: {nth} thread(s) {res[nth][0]:.3e} particle-steps/s ({nsteps} steps, {dt:.1f} s)")
    v, nsteps, dt = res[threads]
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "host": {"nproc": os.cpu_count(), "cpu_model": model, "threads_used": threads},
            "sample": f"{nsteps} Euler steps of a {n_sample}-particle bulk drawn and relaxed like the 1M workload "
                      f"(oracle -O3 -ffast-math, COM link list, race-free gather over {threads} threads) in {dt:.1f} s",
            "value_1core": res[1][0]}


MLP_FLOP_PER_PARTICLE = 2 * (72 * 128 + 128 * 128 + 128 * 3)  # 51,968 useful FLOP per inference
MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_leg(md, ts, n, plain_value, args, p, box, col):
    """Train the drift surrogate on the GPU from plain trajectories of the same state, then
    time the surrogate-assisted step (judge every 16th macro-step, fallback above tol)."""
    cur = md.download(col.copy())  # the state the network is trained and evaluated on
    floor, drift = surrogate.noise_floor(p, box, cur, ts)
    order_floor = surrogate.noise_order_floor(p, box, cur, ts)
    base_err = surrogate.base_predictor_errors(p, box, cur, ts)
    log(f"surrogate: network-free predictors on the same state: {', '.join(f'{k} {v:.3e}' for k, v in base_err.items())}")
    log(f"surrogate: drift's dependence on the order of the 8 increments (floor of any predictor given the "
        f"start state and the noise sum) {order_floor:.3e}")
    log(f"surrogate: deterministic 8-step displacement rms {drift:.3e}; spread over noise (floor of any "
        f"start-state predictor) {floor:.3e}")
    t0 = time.perf_counter()
    ts, losses = surrogate.train(md, ts, 2 << 20, 16, args.train_steps, batch=8192, lr=1e-3, log=log)
    t_train = time.perf_counter() - t0
    nsamp = 2 << 20
    log(f"surrogate: collection + {args.train_steps} Adam steps in {t_train:.1f} s, loss {losses[0]:.3e} -> "
        f"{losses[-50:].mean():.3e}")
    ts, rms, mx = md.surrogate_eval(ts)
    log(f"surrogate: 8-step error rms/max network {rms[0]:.3e}/{mx[0]:.3e} zero-drift {rms[1]:.3e}/{mx[1]:.3e} "
        f"euler-8dt {rms[2]:.3e}/{mx[2]:.3e}")
    nmac = max(args.judge_every, (args.steps + 7) // 8)
    err = None
    kus, st, value = {}, {}, None
    try:
        # per-kernel averages (timing mode synchronises after every macro-step)
        md.set_timing(True)
        ts, _ = md.surrogate_step(args.judge_every, ts, judge_every=args.judge_every, tol=args.tol)
        kus = md.timings()
        md.set_timing(False)
        md.surrogate_stats_reset()
        md.num_cells()
        t0 = time.perf_counter()
        ts, st = md.surrogate_step(nmac, ts, judge_every=args.judge_every, tol=args.tol)
        md.num_cells()
        el = time.perf_counter() - t0
        value = n * 8 * nmac / el
        log(f"surrogate: {nmac} macro-steps in {el * 1e3:.1f} ms -> {value:.3e} particle-steps/s "
            f"({value / plain_value:.2f}x plain); stats {st}")
    except abi.CilError as e:
        err = str(e)
        log(f"surrogate: run failed: {err}")
    mlp_us = kus.get("mlp", 0.0)
    mfma_tf = MLP_FLOP_PER_PARTICLE * n / (mlp_us * 1e-6) / 1e12 if mlp_us > 0 else 0.0
    return {
        "value": value, "unit": "particle-steps/s", "speedup_vs_plain": value / plain_value if value else None,
        "error": err,
        "macro_steps": nmac, "judge_every": args.judge_every, "tol_sigma": args.tol, "stats": st,
        "eval_err_rms": {"network": rms[0], "zero_drift": rms[1], "euler_8dt": rms[2]},
        "eval_err_max": {"network": mx[0], "zero_drift": mx[1], "euler_8dt": mx[2]},
        "noise_floor_rms": floor, "noise_order_floor_rms": order_floor, "drift_rms": drift,
        "base_predictors_rms": base_err,
        "training": {"samples": nsamp, "adam_steps": args.train_steps, "batch": 8192, "seconds": t_train,
                     "loss_first": float(losses[0]), "loss_last50": float(losses[-50:].mean())},
        "kernels_us": {k: kus[k] for k in ("sort", "knn", "mlp", "macro") if k in kus},
        "mfma": {"kernel": "k_mlp_forward", "achieved": mfma_tf, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": mfma_tf / MFMA_BF16_PEAK_TFLOPS, "flop_per_particle": MLP_FLOP_PER_PARTICLE},
    }


def copy_calibration(nbytes=256 << 20, reps=3):
    """Device-to-device copies of a known size: the FETCH_SIZE / WRITE_SIZE counters of the
    runtime's copy kernel calibrate the counters of the library's kernels."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=N_PARTICLES)
    ap.add_argument("--bins-per-cutoff", type=int, default=0,
                    help="pin the bins per cut-off (skips the autotuner)")
    ap.add_argument("--block", type=int, default=0, help="pin the pair kernel block size (with --bins-per-cutoff)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-list", action="store_true", help="stencil pair kernel instead of the neighbour list")
    ap.add_argument("--skin", type=float, default=0.6, help="neighbour-list skin (the outer list's, with --inner-skin)")
    ap.add_argument("--inner-skin", type=float, default=0.2, help="two-level lists: the inner list's skin (0: one level)")
    ap.add_argument("--surrogate-skin", type=float, default=0.3,
                    help="skin of the one-level lists the surrogate legs use")
    ap.add_argument("--repeats", type=int, default=5,
                    help="steady-state repeats of at least one second each (median reported beside value)")
    ap.add_argument("--list-max", type=int, default=256)
    ap.add_argument("--graph-steps", type=int, default=0,
                    help="steps of the graph-path reference window (default: one steady-state repeat)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--judge-every", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 256 MiB device-to-device copies (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {args.gpus}); sharding a colony across GPUs "
              "is designed but not implemented", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    p, box, col = relaxed_state(args.n)
    log(f"relaxed {args.n} particles in {time.perf_counter() - t0:.1f} s")
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    if args.bins_per_cutoff:
        md.set_tuning(args.bins_per_cutoff, args.block)
    md.upload(col)
    md.forces(want=False)
    if not args.no_list:
        md.set_neighbour_list(True, args.skin, args.list_max, inner_skin=args.inner_skin)
        log(f"neighbour list on: skin {args.skin} sigma, inner skin {args.inner_skin} sigma, "
            f"{args.list_max} entries per particle")
    tuned = None
    if not args.bins_per_cutoff:  # stencil: bins x block x kernel; list mode: bins per cut-off
        tuned = md.autotune(reps=3)
        log(f"autotune: {tuned}")
    ts = md.step(args.warmup, 0)
    # per-family breakdown (every family timed; its event records cost GPU time)
    md.set_timing(True)
    ts = md.step(128, ts)
    fam = md.timings()
    # HIP events around the pair kernel only, on the library's stream, from here on
    md.set_timing(True, families=("pair",))
    md.num_cells()  # synchronise
    log(f"warm-up done ({args.warmup} steps); timing {args.steps} steps")
    t0 = time.perf_counter()
    ts = md.step(args.steps, ts)
    md.num_cells()
    elapsed = time.perf_counter() - t0
    timed_pair_us = md.timings()["pair"]
    log(f"timed region: {elapsed * 1e3 / args.steps:.3f} ms/step")
    # steady state (the metric): the median of repeats of at least one second each.  The
    # steps per repeat come from a warm calibration window of at least one second; every
    # repeat is checked to last one second or is run again longer.  The pair kernel's
    # average over the same repeats (list rebuilds included) feeds the roofline.
    per = max(args.steps, 64)
    while True:
        md.num_cells()
        t0 = time.perf_counter()
        ts = md.step(per, ts)
        md.num_cells()
        cal = time.perf_counter() - t0
        if cal >= 1.0:
            break
        per = int(per * 1.25 / max(cal, 1e-3)) + 1
    per = int(per * 1.1 / cal) + 1
    steady, pair_sum = [], 0.0
    while len(steady) < max(args.repeats, 1):
        md.num_cells()
        t0 = time.perf_counter()
        ts = md.step(per, ts)
        md.num_cells()
        el = time.perf_counter() - t0
        if el < 1.0:  # shorter than the metric's minimum: lengthen and run this repeat again
            per = int(per * 1.1 / el) + 1
            continue
        steady.append((per, el))
        pair_sum += md.timings()["pair"] * per
    steady_ms = sorted(e / k for k, e in steady)[len(steady) // 2] * 1e3
    pair_us = pair_sum / sum(k for k, _ in steady)
    log(f"steady state: median {steady_ms:.4f} ms/step over {len(steady)} repeats of {per} steps "
        f"({', '.join(f'{e:.2f}' for _, e in steady)} s); pair kernel {pair_us:.1f} us per launch")
    nb, w = md.bin_info()
    md.set_timing(False)
    # the graph-captured path (no events) over one window of at least one second, for reference
    gsteps = args.graph_steps or per
    md.num_cells()
    t0 = time.perf_counter()
    ts = md.step(gsteps, ts)
    md.num_cells()
    graph_elapsed = time.perf_counter() - t0
    log(f"graph path: {graph_elapsed * 1e3 / gsteps:.4f} ms/step")
    n = args.n
    value = n / (steady_ms * 1e-3)
    fam["pair"] = pair_us
    achieved = PAIR_BYTES_PER_PARTICLE * n / (pair_us * 1e-6) / 1e9 if pair_us > 0 else 0.0
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", ("pair32" if args.no_list else "pair_list") + "_1M_hbm_traffic.json")
    if n == N_PARTICLES and os.path.exists(tpath):  # FETCH/WRITE counter runs, copy-calibrated
        tj = json.load(open(tpath))
        traffic, traffic_src = tj["traffic_bytes_per_launch"], tj.get("source")
    pipeline = None
    ppath = os.path.join(ROOT, "profiles", "pair_list_1M_pipeline.json")
    if n == N_PARTICLES and not args.no_list and os.path.exists(ppath):  # SQ / TA / TD counter runs
        pipeline = json.load(open(ppath))
    out = {
        "metric": "particle-steps/s (plain FP32 step, 3D bulk, 1M particles)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": steady_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 forces / fp64 positions",
        "data": "synthetic (seeded lognormal bulk, 100 push-off relaxation steps)",
        "config": {"workload": "3D periodic bulk, 1048576 single discs, phi=0.6, lognormal sigma CV 0.1, "
                               "LJ 2.5 sigma (species 0 pairs) + soft core, dt=1e-4, kT=0.05",
                   "particles": n, "bins": nb, "bin_width": w, "autotuned": tuned,
                   "neighbour_list": None if args.no_list else {"skin": args.skin, "inner_skin": args.inner_skin,
                                                                 "max_neighbours": args.list_max}},
        "kernels_us": fam,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_pair32<3>" if args.no_list else "k_pair_list<3>",
                     "kernel_us": pair_us, "kernel_us_window": "steady-state repeats (rebuild passes included)",
                     "bytes_per_particle": PAIR_BYTES_PER_PARTICLE,
                     "pipeline": pipeline},
        "step_hbm_gbs": STEP_BYTES_PER_PARTICLE * value / 1e9,
        "value_definition": "median over the steady-state repeats (each >= 1 s) of particles / (seconds per step); "
                            "steps launched directly in 64-step chunks with HIP events around the pair kernel",
        "steady_state": {
            "repeats": len(steady), "steps_each": [k for k, _ in steady],
            "seconds_each": [round(e, 3) for _, e in steady],
            "median_ms_per_step": steady_ms, "calibration_seconds": round(cal, 3)},
        "timed_region": {"steps": args.steps, "ms_per_step": elapsed * 1e3 / args.steps,
                         "value": n * args.steps / elapsed, "pair_us": timed_pair_us,
                         "note": "the K steps of the command line; kernels_us other than pair from a separate "
                                 "128-step pass timing every family"},
        "graph_path": {"steps": gsteps, "ms_per_step": graph_elapsed * 1e3 / gsteps, "value": n * gsteps / graph_elapsed},
    }
    if not args.no_surrogate:
        if not args.no_list:  # the surrogate runs on one-level lists
            md.set_neighbour_list(True, args.surrogate_skin, args.list_max)
        out["surrogate"] = surrogate_leg(md, ts, n, value, args, p, box, col)
    md.close()
    if not args.no_surrogate:
        # the same relaxed state at kT = 0.005, where the start-state noise floor is below tol
        q = abi.Params.from_buffer_copy(p)
        for s_ in range(q.nspecies):
            q.species[s_].kT = 0.005
        m2 = HipMD(q, box, [sp["capacity"] for sp in col.species])
        # the same plain path as the main leg for the fallback steps
        if tuned:
            m2.set_tuning(tuned["bins_per_cutoff"], tuned["block"])
        elif args.bins_per_cutoff:
            m2.set_tuning(args.bins_per_cutoff, args.block)
        m2.upload(col)
        if not args.no_list:
            m2.set_neighbour_list(True, args.surrogate_skin, args.list_max)
        m2.forces(want=False)
        res = surrogate_leg(m2, 0, n, value, args, q, box, col)
        res["kT"] = 0.005
        out["surrogate_low_kT"] = res
        m2.close()
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # the baseline must not hide the GPU number
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
