"""Benchmark of the MI355X simulation step on the 1M-neuron AdEx network.

Workload (config 3): 1,000,000 aeif_cond_beta_multisynapse neurons, 2
receptor ports (E_rev 0/-85 mV, 1 ms beta time constants), 80/20 exc/inh,
fixed indegree 1000 (800 exc -> port 0, 200 inh -> port 1; 1e9 synapses),
weights U(0.025,0.075) nS exc / U(0.175,0.525) nS inh, delays U{1..50} x 0.1
ms, one Poisson generator (20 kHz, 0.237 nS) per neuron tuned to ~5 Hz,
FP32 state, FP64 time, dt = 0.1 ms.  A "step" is one 0.1 ms simulation step
of the whole network.  The network is built through the nestgpu API, as a
user script would; build and calibration are not timed (as the reference's
"Time to simulate").

Prints ONE JSON line on stdout.  Progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nest-gpu_amd"))
sys.path.insert(0, ROOT)

DT = 0.1
N_FULL = 1_000_000
K_EXC, K_INH = 800, 200
POISSON_RATE, POISSON_W = 20000.0, 0.237
RECEPTORS = {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0], "tau_decay": [1.0, 1.0]}
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rk5_flops_per_attempt(n_port):
    """Algorithmic FP32 flops of one RK5 attempt of aeif_cond_beta_multisynapse:
    stage/error arithmetic 56*NVAR plus 6 right-hand sides of 15+6P flops
    (the exponential counted as one)."""
    nv = 2 + 2 * n_port
    return 56 * nv + 6 * (15 + 6 * n_port)


def build_network(ngpu, n, seed):
    rng = np.random.default_rng(seed)
    ne = int(0.8 * n)
    neurons = ngpu.Create("aeif_cond_beta_multisynapse", n, 2)
    ngpu.SetStatus(neurons, RECEPTORS)
    pg = ngpu.Create("poisson_generator")
    ngpu.SetStatus(pg, "rate", POISSON_RATE)
    ngpu.Connect(pg, neurons, {"rule": "all_to_all"},
                 {"weight": POISSON_W, "delay": DT, "receptor": 0})
    lib = ngpu._nestgpu
    for (src, k, port, lo, hi) in ((neurons[0:ne], K_EXC, 0, 0.025, 0.075),
                                   (neurons[ne:n], K_INH, 1, 0.175, 0.525)):
        m = n * k
        w = rng.random(m, dtype=np.float32)
        w *= np.float32(hi - lo)
        w += np.float32(lo)
        d = rng.integers(1, 51, size=m, dtype=np.int16).astype(np.float32)
        d *= np.float32(DT)
        ngpu.ConnSpecInit()
        ngpu.SynSpecInit()
        ngpu.SetConnSpecParam("rule", 3)
        ngpu.SetConnSpecParam("indegree", k)
        ngpu.SetSynSpecIntParam("receptor", port)
        lib.NESTGPU_SetSynSpecFloatPtParam(b"weight_array", w.ctypes.data_as(ctypes.c_void_p))
        lib.NESTGPU_SetSynSpecFloatPtParam(b"delay_array", d.ctypes.data_as(ctypes.c_void_p))
        lib.NESTGPU_ConnectSeqSeq(src.i0, src.n, neurons.i0, neurons.n)
        ngpu._check()
        del w, d
    return neurons


def cpu_baseline(n_sample, sim_ms, seed):
    """The oracle (host C++/OpenMP restatement of the reference step) on a
    slice of the same network: n_sample neurons with the full in-degree 1000
    and the same per-neuron input; the per-neuron-step cost is scaled to the
    1M-neuron workload."""
    from oracle.oracle import Oracle, threads
    rng = np.random.default_rng(seed)
    n = n_sample
    ne = int(0.8 * n)
    o = Oracle(prec=0)
    n0 = o.create("aeif_cond_beta_multisynapse", n, 2)
    for p in range(2):
        o.set(n0, n, "E_rev", [RECEPTORS["E_rev"][p]] * n, port=p)
        o.set(n0, n, "tau_rise", [1.0] * n, port=p)
        o.set(n0, n, "tau_decay", [1.0] * n, port=p)
    pg = o.create("poisson_generator", 1, 0)
    o.set(pg, 1, "rate", [POISSON_RATE])
    o.connect(np.full(n, pg, np.int32), np.arange(n, dtype=np.int32) + n0, 0, POISSON_W, DT)
    for (lo_src, hi_src, k, port, lo, hi) in ((0, ne, K_EXC, 0, 0.025, 0.075),
                                               (ne, n, K_INH, 1, 0.175, 0.525)):
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
    per_neuron_step = wall / (n * steps)
    rtf_full = per_neuron_step * N_FULL * (1000.0 / DT)
    return {"value": rtf_full, "unit": "s_wall/s_bio", "cores": threads(), "kind": "port",
            "sample": "oracle on %d neurons x %d synapses/neuron for %.0f ms (%.1f s wall); "
                      "per-neuron-step cost scaled to 1e6 neurons x 1 s" % (n, K_EXC + K_INH, sim_ms, wall),
            "neuron_updates_per_s": n * steps / wall,
            "syn_events_per_s": (st1["syn_events"] - st0["syn_events"]) / wall}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000, help="timed 0.1 ms steps (10000 = 1 s)")
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--neurons", type=int, default=N_FULL)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=10000)
    ap.add_argument("--cpu-ms", type=float, default=100.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB fill and a 1 GiB elementwise multiply (torch) whose "
                         "FETCH_SIZE/WRITE_SIZE counters calibrate the traffic figures")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (multi-GPU spike exchange is out of "
              "scope); run with --gpus 1", file=sys.stderr)
        sys.exit(0)

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
    ngpu.SetRandomSeed(args.seed)
    n = args.neurons
    t0 = time.perf_counter()
    build_network(ngpu, n, args.seed)
    log("built %d neurons, %d connections in %.1f s" % (n, n * (K_EXC + K_INH) + n, time.perf_counter() - t0))
    t0 = time.perf_counter()
    ngpu.Calibrate()
    log("calibrated in %.1f s" % (time.perf_counter() - t0))
    ngpu.Simulate(args.warmup * DT)
    ngpu.Synchronize()
    log("warmup %d steps done" % args.warmup)
    st0 = ngpu.GetStepStats()
    ngpu.SetKernelTiming(True)
    ngpu.Synchronize()
    t0 = time.perf_counter()
    ngpu.Simulate(args.steps * DT)
    ngpu.Synchronize()
    wall = time.perf_counter() - t0
    kt = ngpu.GetKernelTimes()
    st1 = ngpu.GetStepStats()
    ngpu.SetKernelTiming(False)
    bio_s = args.steps * DT / 1000.0
    rtf = wall / bio_s
    att = st1["attempts"] - st0["attempts"]
    rej = st1["rejected"] - st0["rejected"]
    ev = st1["syn_events"] - st0["syn_events"]
    spk = st1["spikes"] - st0["spikes"]
    log("timed %d steps: wall %.3f s, RTF %.4f, kernel ms %s" % (args.steps, wall, rtf, kt))

    # dominant kernel and its roofline
    dom = max(kt, key=lambda k: kt[k][0])
    rk5_ms, rk5_n = kt["rk5_update"]
    if dom == "rk5_update":
        flops_total = att * rk5_flops_per_attempt(2)
        achieved = flops_total / (rk5_ms * 1e-3) / 1e12
        roof = {"kernel": "rk5_kernel<AeifCondBeta<2>>", "bound": "mfma", "unit": "TFLOP/s",
                "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "frac": achieved / PEAK_FP32_TFLOPS,
                "traffic": None, "avg_launch_us": rk5_ms * 1e3 / max(rk5_n, 1),
                "flops_per_launch": flops_total / max(rk5_n, 1)}
    else:
        ms, cnt = kt[dom]
        bytes_total = ev * 10.0 if dom == "deliver" else n * 2 * 16.0 * cnt
        achieved = bytes_total / (ms * 1e-3) / 1e9
        roof = {"kernel": dom, "bound": "hbm", "unit": "GB/s", "achieved": achieved,
                "peak": PEAK_HBM_GBS, "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                "avg_launch_us": ms * 1e3 / max(cnt, 1)}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: oracle on %d neurons for %.0f ms" % (args.cpu_sample, args.cpu_ms))
        cpu = cpu_baseline(args.cpu_sample, args.cpu_ms, args.seed)

    out = {
        "metric": "real-time factor (1M AdEx, 1e9 syn) + synaptic events/s; learned-h vs plain RK5",
        "value": rtf,
        "unit": "s_wall/s_bio",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "aeif_cond_beta_multisynapse %d neurons, 2 ports, indegree 1000 "
                               "(%.2g synapses), delays U{1..50}x0.1 ms, Poisson 20 kHz, plain RK5, "
                               "1 s bio" % (n, n * 1000.0),
                   "neurons": n, "synapses": n * 1000, "dt_ms": DT, "method": "plain RK5 (eps 1e-6)"},
        "synaptic_events_per_s": ev / wall,
        "neuron_updates_per_s": n * args.steps / wall,
        "mean_rate_hz": spk / n / bio_s,
        "dropped_spikes": st1["dropped_spikes"],
        "rk5_attempts_per_neuron_step": att / (n * args.steps),
        "rk5_rejected_fraction": rej / max(att, 1),
        "kernel_ms": {k: v[0] for k, v in kt.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
