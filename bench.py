"""Benchmark of the MI355X simulation step on the 1M-neuron AdEx networks.

Default workload (scale_adex, nest-gpu_amd/workloads.py): 1,000,000
aeif_cond_beta_multisynapse neurons, 2 receptor ports, 80/20 exc/inh, fixed
indegree 1000 (1e9 synapses), delays U{1..50} x 0.1 ms, Poisson drive tuned to
~5 Hz, FP32 state, FP64 time, dt = 0.1 ms.  A "step" is one 0.1 ms simulation
step of the whole network.  The network is built through the nestgpu API, as a
user script would; build and calibration are not timed (as the reference's
"Time to simulate").

Timing: at least 3000 untimed steps (300 ms of biological time, in which
the network's rate settles) whatever --warmup asks, then exactly `--steps`
timed steps with the plain RK5 controller (the reported `value`; a window
that delivers no synaptic event is refused).  When a trained step-size
predictor exists for the workload (nest-gpu_amd/predictors/<workload>.json or
--predictor), three further 1 s windows follow on the same network: plain,
learned, plain again, with the learned window's RK5 statistics, spike rate
and ISI CV against the mean of the two plain windows around it.

Prints ONE JSON line on stdout.  Progress goes to stderr.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nest-gpu_amd"))
sys.path.insert(0, ROOT)

import workloads  # noqa: E402

DT = workloads.DT
N_FULL = 1_000_000
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# untimed steps before the first timed window: the network's rate settles
# within ~300 ms of biological time
SETTLE_STEPS = 3000
# steps of each window of the learned-vs-plain comparison (1 s bio)
LEARN_STEPS = 10000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rk5_flops_per_attempt(n_port):
    """Algorithmic FP32 flops of one RK5 attempt of aeif_cond_beta_multisynapse:
    stage/error arithmetic 56*NVAR plus 6 right-hand sides of 15+6P flops
    (the exponential counted as one)."""
    nv = 2 + 2 * n_port
    return 56 * nv + 6 * (15 + 6 * n_port)


# algorithmic bytes: delivery per (spike, synapse) event = synapse record
# (u32 channel + f32 weight + u16 delay) + FP64 accumulator read and write
DELIVER_BYTES_PER_EVENT = 4 + 4 + 2 + 16
# fold per (neuron, port) channel = ring slot read + clear (16 B) + input
# variable read and write (8 B); Poisson edges kept as per-channel CSR add
# 20 B each, edges of a uniform run (all-to-all or one-to-one from
# generators, as in every workload here) none: weight, delay and stream are
# launch arguments and the generator's table entry stays in cache
FOLD_BYTES_PER_CHANNEL = 24
# dense (tiled) delivery: an 8-byte record per event, the (first, end)
# offsets of every (spike, tile) pair, and one FP64 ring read and write per
# ring entry that received input (at most one per event, at most every
# entry of the LDS rows per launch)
TILE_REC_BYTES = 8
TILE_OFF_BYTES = 8
RING_RMW_BYTES = 16
# iaf_psc_exp update: 4 state floats read and written per neuron-step
IAF_BYTES_PER_NEURON = 32
POISSON_BYTES_PER_EDGE = 20


def cpu_model():
    """The host CPU's model line (lscpu 'Model name', else /proc/cpuinfo)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.strip().startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_slice(n_sample, sim_ms, seed):
    """The oracle (host C++/OpenMP restatement of the reference step) on a
    slice of the scale_adex network: n_sample neurons with the full in-degree
    1000 and the same per-neuron input.  Returns (wall seconds per
    neuron-step, synaptic events per wall second, wall seconds)."""
    from oracle.oracle import Oracle
    rng = np.random.default_rng(seed)
    n = n_sample
    ne = int(0.8 * n)
    o = Oracle(prec=0)
    n0 = o.create("aeif_cond_beta_multisynapse", n, 2)
    for p in range(2):
        for k, v in workloads.SCALE_RECEPTORS.items():
            o.set(n0, n, k, [v[p]] * n, port=p)
    pg = o.create("poisson_generator", 1, 0)
    o.set(pg, 1, "rate", [workloads.SCALE_POISSON[0]])
    o.connect(np.full(n, pg, np.int32), np.arange(n, dtype=np.int32) + n0, 0,
              workloads.SCALE_POISSON[1], DT)
    for (lo_src, hi_src, k, port, lo, hi) in ((0, ne, 800, 0, 0.025, 0.075),
                                               (ne, n, 200, 1, 0.175, 0.525)):
        m = n * k
        src = rng.integers(lo_src, hi_src, size=m).astype(np.int32) + n0
        tgt = np.repeat(np.arange(n, dtype=np.int32), k) + n0
        w = rng.uniform(lo, hi, size=m).astype(np.float32)
        d = (rng.integers(1, 51, size=m) * DT).astype(np.float32)
        o.connect(src, tgt, port, w, d)
    o.simulate(20.0)   # calibration and transient, untimed
    st0 = o.stats()
    t0 = time.perf_counter()
    o.simulate(sim_ms)
    wall = time.perf_counter() - t0
    st1 = o.stats()
    steps = int(round(sim_ms / DT))
    return wall / (n * steps), (st1["syn_events"] - st0["syn_events"]) / wall, wall


def cpu_baseline(n_sample, sim_ms, seed):
    """CPU baseline of the 1M-neuron second: the oracle's per-neuron-step
    cost on an n_sample slice, scaled to 1e6 neurons x 1 s.  A second slice
    of half the size states how far the per-neuron-step cost moves with the
    slice size (the extrapolation's error estimate: the full network's state
    and synapse table no longer fit the host caches, so the scaled figure is
    a lower bound of the full-size CPU time)."""
    from oracle.oracle import threads
    half = max(n_sample // 2, 1000)
    c_half, _, w_half = cpu_slice(half, sim_ms, seed)
    c_full, ev, wall = cpu_slice(n_sample, sim_ms, seed)
    rtf_full = c_full * N_FULL * (1000.0 / DT)
    return {"value": rtf_full, "unit": "s_wall/s_bio", "cores": threads(), "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "oracle on %d neurons x %d synapses/neuron for %.0f ms (%.1f s wall); "
                      "per-neuron-step cost scaled to 1e6 neurons x 1 s" % (n_sample, 1000, sim_ms, wall),
            "extrapolation_check": {"neurons": half, "wall_s": w_half,
                                    "cost_ratio_full_vs_half": c_full / c_half,
                                    "note": "per-neuron-step cost of the %d-neuron slice over that "
                                            "of the %d-neuron slice" % (n_sample, half)},
            "neuron_updates_per_s": 1.0 / c_full,
            "syn_events_per_s": ev}


# sampled kernel timing (modes 3 and 4): one step in TIMING_STRIDE is
# bracketed by events
TIMING_STRIDE = 16


def timed_window(ngpu, steps, timing=2):
    """Simulate `steps` steps between two synchronisations; returns wall time,
    kernel times (HIP events on the simulator stream: timing 1 brackets every
    launch, 2 only the RK5 launches, 3 / 4 the same in one step of every
    TIMING_STRIDE, scaled up here to the whole window) and statistics deltas."""
    st0 = ngpu.GetStepStats()
    ngpu.SetKernelTiming(timing)
    ngpu.Synchronize()
    t0 = time.perf_counter()
    ngpu.Simulate(steps * DT)
    ngpu.Synchronize()
    wall = time.perf_counter() - t0
    kt = ngpu.GetKernelTimes()
    if timing in (3, 4):
        kt = {k: (ms * TIMING_STRIDE, cnt * TIMING_STRIDE) for k, (ms, cnt) in kt.items()}
    st1 = ngpu.GetStepStats()
    ngpu.SetKernelTiming(False)
    d = {k: st1[k] - st0[k] for k in ("attempts", "rejected", "substeps", "syn_events", "spikes",
                                      "hot_listed", "split_steps")}
    return {"wall": wall, "kt": kt, "stats": d}


def window_spike_stats(times, t0_ms, t1_ms, n):
    """Mean rate over all neurons and mean ISI CV over neurons with >= 4
    spikes inside [t0, t1)."""
    cnt = 0
    cvs = []
    for ts in times:
        a = np.asarray(ts, np.float64)
        a = a[(a >= t0_ms) & (a < t1_ms)]
        cnt += a.size
        if a.size >= 4:
            isi = np.diff(np.sort(a))
            cvs.append(isi.std() / isi.mean())
    return cnt / n / ((t1_ms - t0_ms) * 1e-3), (float(np.mean(cvs)) if cvs else None), len(cvs)


def rooflines(w, n, n_port, steps, n_poisson_edges, tiled=None):
    kt, st = w["kt"], w["stats"]
    out = {}
    ms, cnt = kt["rk5_update"]
    if cnt:
        flops = st["attempts"] * rk5_flops_per_attempt(n_port)
        a = flops / (ms * 1e-3) / 1e12
        out["rk5_update"] = {"bound": "fp32_valu", "unit": "TFLOP/s", "achieved": a, "peak": PEAK_FP32_TFLOPS,
                             "frac": a / PEAK_FP32_TFLOPS, "avg_launch_us": ms * 1e3 / cnt,
                             "flops_per_launch": flops / cnt,
                             "note": "FP32 vector peak (the RK5 arithmetic is FP32 VALU work)"}
    ms, cnt = kt["iaf_update"]
    if cnt:
        b = n * IAF_BYTES_PER_NEURON * steps
        a = b / (ms * 1e-3) / 1e9
        out["iaf_update"] = {"bound": "hbm", "unit": "GB/s", "achieved": a, "peak": PEAK_HBM_GBS,
                             "frac": a / PEAK_HBM_GBS, "avg_launch_us": ms * 1e3 / cnt,
                             "bytes_per_launch": b / cnt}
    ms, cnt = kt["deliver"]
    if cnt and cnt < steps // 2:
        out["deliver"] = {"note": "delivery runs inside the update launches (fused); the %d "
                                  "delivery launch(es) only close Simulate calls" % cnt}
    elif cnt and tiled:
        ev = st["syn_events"]
        entries = n * n_port * tiled["tile_rows"]
        b = (ev * TILE_REC_BYTES + st["spikes"] * tiled["n_tiles"] * TILE_OFF_BYTES
             + RING_RMW_BYTES * min(ev, cnt * entries))
        a = b / (ms * 1e-3) / 1e9
        out["deliver"] = {"bound": "hbm", "unit": "GB/s", "achieved": a, "peak": PEAK_HBM_GBS,
                          "frac": a / PEAK_HBM_GBS, "avg_launch_us": ms * 1e3 / cnt,
                          "bytes_per_launch": b / cnt,
                          "note": "dense tiled delivery: 8 B record per event + 8 B offsets per "
                                  "(spike, tile) + 16 B per ring entry with input"}
    elif cnt:
        b = st["syn_events"] * DELIVER_BYTES_PER_EVENT
        a = b / (ms * 1e-3) / 1e9
        out["deliver"] = {"bound": "hbm", "unit": "GB/s", "achieved": a, "peak": PEAK_HBM_GBS,
                          "frac": a / PEAK_HBM_GBS, "avg_launch_us": ms * 1e3 / cnt,
                          "bytes_per_launch": b / cnt}
    ms, cnt = kt["fold_poisson"]
    if cnt:
        # one launch per population with ports and step
        b = (n * n_port * FOLD_BYTES_PER_CHANNEL + n_poisson_edges * POISSON_BYTES_PER_EDGE) * cnt
        a = b / (ms * 1e-3) / 1e9
        out["fold_poisson"] = {"bound": "hbm", "unit": "GB/s", "achieved": a, "peak": PEAK_HBM_GBS,
                               "frac": a / PEAK_HBM_GBS, "avg_launch_us": ms * 1e3 / cnt,
                               "bytes_per_launch": b / cnt}
    return out


def window_summary(w, steps, n):
    bio = steps * DT / 1000.0
    st = w["stats"]
    return {"rtf": w["wall"] / bio, "ms_per_step": w["wall"] * 1e3 / steps,
            "neuron_updates_per_s": n * steps / w["wall"],
            "synaptic_events_per_s": st["syn_events"] / w["wall"],
            "rk5_attempts_per_neuron_step": st["attempts"] / (n * steps),
            "rk5_rejected_fraction": st["rejected"] / max(st["attempts"], 1),
            "mean_rate_hz": st["spikes"] / n / bio,
            "hot_listed_per_step": st["hot_listed"] / steps,
            "split_steps_per_step": st["split_steps"] / steps,
            "kernel_ms": {k: v[0] for k, v in w["kt"].items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000, help="timed 0.1 ms steps per window (10000 = 1 s)")
    ap.add_argument("--warmup", type=int, default=SETTLE_STEPS,
                    help="untimed steps first; at least %d (300 ms, the network's rate settles "
                         "within that) whatever is asked" % SETTLE_STEPS)
    ap.add_argument("--workload", default="scale_adex", choices=sorted(workloads.WORKLOADS))
    ap.add_argument("--neurons", type=int, default=N_FULL)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--predictor", default=None,
                    help="trained predictor JSON ('none' to skip the learned window)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=20000)
    ap.add_argument("--cpu-ms", type=float, default=1000.0)
    ap.add_argument("--tune-steps", type=int, default=20,
                    help="batching workload: steps per autotuner candidate (0: no autotuning)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB fill and a 1 GiB elementwise multiply (torch) whose "
                         "FETCH_SIZE/WRITE_SIZE counters calibrate the traffic figures")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (multi-GPU spike exchange is out of "
              "scope); run with --gpus 1", file=sys.stderr)
        sys.exit(0)
    mc = args.workload == "microcircuit"
    bt = args.workload == "batching"
    seed = args.seed if args.seed is not None else {"scale_adex": 12345, "microcircuit": 7,
                                                    "batching": 300}.get(args.workload, 200)
    pred_path = args.predictor
    if pred_path is None:
        p = os.path.join(ROOT, "nest-gpu_amd", "predictors", args.workload + ".json")
        pred_path = p if os.path.exists(p) else "none"
    predictor = None
    if pred_path != "none":
        with open(pred_path) as f:
            predictor = json.load(f)

    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")   # fill: writes 1 GiB
        y = torch.empty_like(x)
        torch.mul(x, 2.0, out=y)                                       # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        log("copy calibration: fill 1 GiB, then elementwise 1 GiB read + 1 GiB written")
        del x, y

    import nestgpu as ngpu
    ngpu.SetVerbosityLevel(0)
    ngpu.SetStepPredictorMode("plain")
    n = (sum(workloads.MC_SIZES) if mc else
         sum(p[1] for p in workloads.BATCH_POPS) if bt else args.neurons)
    n_port = 4 if args.workload == "predictor_adex" else 2
    n_poisson_edges = n          # Poisson edges built (one per neuron)
    n_poisson_csr_edges = 0      # of which kept as CSR (each workload's are one run per population)
    t0 = time.perf_counter()
    neurons = workloads.WORKLOADS[args.workload](ngpu, n, seed)
    n_syn = (workloads.microcircuit.n_syn if mc else workloads.batching.n_syn if bt else
             n * 1000)
    log("built %s: %d neurons, %d synapses + %d Poisson edges in %.1f s"
        % (args.workload, n, n_syn, n, time.perf_counter() - t0))
    # events around a launch add a gap to its step: the timed windows bracket
    # the launches of one step in TIMING_STRIDE only (the microcircuit, which
    # has no RK5, every launch of those steps; the others their RK5 launch)
    win_timing = 3 if mc else 4
    if predictor is not None:
        ngpu.ActivateRecSpikeTimes(neurons, 400)
    t0 = time.perf_counter()
    ngpu.Calibrate()
    log("calibrated in %.1f s" % (time.perf_counter() - t0))
    # the network's rate settles within ~300 ms of biological time: whatever
    # --warmup asks, at least that much runs untimed before any window
    settle = max(args.warmup, SETTLE_STEPS)
    ngpu.Simulate(settle * DT)
    ngpu.Synchronize()
    log("settled: %d untimed steps (%.0f ms bio)" % (settle, settle * DT))
    t_plain0 = settle * DT
    batch = None
    if bt:
        batch = {"config_default": ngpu.GetBatchConfig()}
        if args.tune_steps > 0:
            t0 = time.perf_counter()
            best_ms, ncand = ngpu.AutotuneBatch(args.tune_steps)
            batch["autotune"] = {"candidates": ncand, "steps_per_candidate": args.tune_steps,
                                 "best_ms_per_step": best_ms,
                                 "wall_s": time.perf_counter() - t0}
            t_plain0 += ncand * args.tune_steps * DT
            log("autotuned the batch kernel: %d candidates x %d steps, best %.4f ms per step"
                % (ncand, args.tune_steps, best_ms))
        batch["config"] = ngpu.GetBatchConfig()
        log("batch configuration: %s" % batch["config"])
    else:
        batch = {"config": ngpu.GetBatchConfig()}
        log("batch configuration: %s" % batch["config"])
    wp = timed_window(ngpu, args.steps, timing=win_timing)
    sp = window_summary(wp, args.steps, n)
    log("plain: %d steps, wall %.3f s, RTF %.4f, events %d, kernel ms %s"
        % (args.steps, wp["wall"], sp["rtf"], wp["stats"]["syn_events"], wp["kt"]))
    if wp["stats"]["syn_events"] == 0:
        # a window without synaptic events is not the workload: no value
        log("bench.py: the timed window delivered no synaptic events (network inactive); "
            "no measurement is reported")
        sys.exit(3)
    t_now = t_plain0 + args.steps * DT

    learned = None
    if predictor is not None:
        # plain, learned, plain again on fixed 1 s windows after the timed
        # one (whatever --steps is): the learned window's spike statistics
        # are compared with the mean of the two plain windows around it, so a
        # slow drift of the network is not read as an effect
        L = LEARN_STEPS
        ngpu.Simulate(200 * DT)     # untimed gap
        t_pa = t_now + 200 * DT
        wa = timed_window(ngpu, L, timing=win_timing)
        sa = window_summary(wa, L, n)
        ngpu.SetPredictorParams(np.asarray(predictor["params"], np.float32))
        ngpu.SetStepPredictorMode("learned")
        ngpu.Simulate(200 * DT)     # switch-over, untimed
        t_learn0 = t_pa + (L + 200) * DT
        wl = timed_window(ngpu, L, timing=win_timing)
        ngpu.SetStepPredictorMode("plain")
        sl = window_summary(wl, L, n)
        log("learned: wall %.3f s, RTF %.4f, kernel ms %s" % (wl["wall"], sl["rtf"], wl["kt"]))
        ngpu.Simulate(200 * DT)
        t_plain2 = t_learn0 + (L + 200) * DT
        wp2 = timed_window(ngpu, L, timing=win_timing)
        sp2 = window_summary(wp2, L, n)
        log("plain before/after: RTF %.4f / %.4f" % (sa["rtf"], sp2["rtf"]))
        t0 = time.perf_counter()
        times = ngpu.GetRecSpikeTimes(neurons)
        rp, cvp, ncp = window_spike_stats(times, t_pa, t_pa + L * DT, n)
        rl, cvl, ncl = window_spike_stats(times, t_learn0, t_learn0 + L * DT, n)
        rp2, cvp2, _ = window_spike_stats(times, t_plain2, t_plain2 + L * DT, n)
        log("spike statistics in %.1f s: plain rate %.3f CV %s, learned rate %.3f CV %s, plain again "
            "rate %.3f CV %s" % (time.perf_counter() - t0, rp, cvp, rl, cvl, rp2, cvp2))
        sa["isi_cv"] = cvp
        sl["isi_cv"] = cvl
        sp2["isi_cv"] = cvp2
        r_ref = 0.5 * (rp + rp2)
        cv_ref = 0.5 * (cvp + cvp2) if (cvp and cvp2) else None
        learned = dict(sl)
        learned["window_steps"] = L
        learned["plain_before"] = sa
        learned["plain_after"] = sp2
        learned["speedup_vs_plain"] = 0.5 * (sa["rtf"] + sp2["rtf"]) / sl["rtf"]
        learned["rk5_kernel_speedup_vs_plain"] = (0.5 * (wa["kt"]["rk5_update"][0] + wp2["kt"]["rk5_update"][0])
                                                  / max(wl["kt"]["rk5_update"][0], 1e-9))
        learned["rate_rel_diff"] = (rl - r_ref) / r_ref if r_ref else None
        learned["cv_rel_diff"] = (cvl - cv_ref) / cv_ref if (cv_ref and cvl) else None
        learned["neurons_in_cv"] = ncl
        learned["predictor"] = {k: predictor[k] for k in predictor if k != "params"}

    # every launch timed (events around each kernel slow the step a little,
    # so this window is reported apart from the timed windows above): the
    # per-kernel split of a step and the rooflines of all kernels
    nb = 2000
    wb = timed_window(ngpu, nb, timing=1)
    log("breakdown: %d steps, wall %.3f s, kernel ms %s" % (nb, wb["wall"], wb["kt"]))
    tiled = batch["config"] if batch and batch["config"].get("tiled") else None
    roof_all = rooflines(wb, n, n_port, nb, n_poisson_csr_edges, tiled)
    breakdown = {k: v[0] * 1e3 / nb for k, v in wb["kt"].items()}
    # the dominant kernel's roofline from the plain timed window (its
    # sampled launches), else from the breakdown window; HBM traffic per
    # launch from
    # the committed FETCH_SIZE/WRITE_SIZE summary of this workload, if any
    dom = max(wb["kt"], key=lambda k: wb["kt"][k][0])
    roof = dict(rooflines(wp, n, n_port, args.steps, n_poisson_csr_edges, tiled).get(dom, roof_all.get(dom, {})))
    rk5_kname = ("rk5_batch_kernel<%d>" % (3 + batch["config"]["variant"]) if bt else
                 "rk5_persistent_kernel<AeifCondBeta<%d>, true, 0>" % n_port)
    kname = {"rk5_update": rk5_kname,
             "deliver": "deliver_tiled_kernel" if tiled else "deliver_kernel",
             "fold_poisson": "fold_kernel<%d>" % n_port,
             "iaf_update": "iaf_kernel"}.get(dom, dom)
    roof["kernel"] = kname
    roof["traffic"] = None
    tpath = os.path.join(ROOT, "profiles", "%s_traffic.json" % args.workload)
    if os.path.exists(tpath) and (n == N_FULL or mc):
        with open(tpath) as f:
            tr = json.load(f)
        if kname in tr["kernels"]:
            roof["traffic"] = tr["kernels"][kname]["bytes"]
            roof["traffic_source"] = "profiles/%s_traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE, " \
                                     "scaled by fill/copy calibration)" % args.workload

    cpu = None
    if not args.no_cpu_baseline and args.workload == "scale_adex":
        log("cpu baseline: oracle on %d neurons for %.0f ms" % (args.cpu_sample, args.cpu_ms))
        cpu = cpu_baseline(args.cpu_sample, args.cpu_ms, seed)

    out = {
        "metric": "real-time factor (1M AdEx, 1e9 syn) + synaptic events/s; learned-h vs plain RK5",
        "value": sp["rtf"],
        "unit": "s_wall/s_bio",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_steps": settle,
        "timed_bio_ms": args.steps * DT,
        "ms_per_step": sp["ms_per_step"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("%s: %d neurons in six populations of five models (300 + 3000 "
                                "izhikevich_cond_beta, 30000 aeif_psc_exp, 120000 "
                                "aeif_cond_alpha_multisynapse with 4 ports, 300000 "
                                "aeif_cond_beta_multisynapse with 2 ports, 50000 iaf_psc_exp), fixed "
                                "indegree 500 across populations (%.3g synapses), normal weights and "
                                "delays, Poisson input; one batch launch per step; timed %d steps = "
                                "%.1f ms bio after %.0f ms bio settling and autotuning"
                                % (args.workload, n, n_syn, args.steps, args.steps * DT,
                                   t_plain0)) if bt else
                               ("%s: iaf_psc_exp %d neurons in 8 populations, fixed_total_number per "
                                "pair (p ~ U(0.01,0.15)), %.3g synapses, normal weights and delays, "
                                "Poisson 8 Hz x K_ext, exact propagator; timed %d steps = %.1f ms bio after "
                                "%.0f ms bio settling" % (args.workload, n, n_syn, args.steps,
                                                          args.steps * DT, settle * DT)) if mc else
                               ("%s: aeif_cond_beta_multisynapse %d neurons, %d ports, indegree 1000 "
                                "(%.2g synapses), delays U{1..50}x0.1 ms, Poisson input, plain RK5 "
                                "(eps 1e-6); timed %d steps = %.1f ms bio after %.0f ms bio settling"
                                % (args.workload, n, n_port, n_syn, args.steps, args.steps * DT,
                                   settle * DT)),
                   "neurons": n, "synapses": n_syn, "dt_ms": DT, "seed": seed,
                   "method": "exact iaf_psc_exp update" if mc else
                             "plain RK5 + exact iaf_psc_exp in one batch launch per step" if bt else
                             "plain RK5 (value); learned step-size proposal in 'learned'"},
        "synaptic_events_per_s": sp["synaptic_events_per_s"],
        "neuron_updates_per_s": sp["neuron_updates_per_s"],
        "mean_rate_hz": sp["mean_rate_hz"],
        "rk5_attempts_per_neuron_step": sp["rk5_attempts_per_neuron_step"],
        "rk5_rejected_fraction": sp["rk5_rejected_fraction"],
        "kernel_ms": sp["kernel_ms"],
        "kernel_us_per_step": breakdown,
        "plain": sp,
        "learned": learned,
        "roofline": roof,
        "rooflines": roof_all,
        "cpu_baseline": cpu,
    }
    out["batch"] = batch
    if bt:
        out["roofline"]["note"] = ("flops counted at aeif_cond_beta_multisynapse (2 ports) per RK5 "
                                   "attempt for every model of the batch; iaf_psc_exp updates not "
                                   "counted")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
