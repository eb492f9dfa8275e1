"""Train the learned RK5 step-size predictor for a workload and save it.

For every seed the workload network is built, simulated briefly with the
plain controller, then simulated with (state, accepted h) logging of every
substep of every `stride`-th neuron; the pooled records train the MLP on the
GPU (the library's own forward/backward and Adam kernels).  The parameter
vector is written as JSON for `bench.py --predictor` or SetPredictorParams.

    python nest-gpu_amd/train_predictor.py --workload predictor_adex \\
        --seeds 100-109 --pairs 1e7 --out nest-gpu_amd/predictors/predictor_adex.json
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import nestgpu as ngpu  # noqa: E402
import workloads  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def seed_list(spec):
    if "-" in spec:
        a, b = spec.split("-")
        return list(range(int(a), int(b) + 1))
    return [int(s) for s in spec.split(",")]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="predictor_adex", choices=sorted(workloads.WORKLOADS))
    ap.add_argument("--neurons", type=int, default=1_000_000)
    ap.add_argument("--seeds", default="100-109")
    ap.add_argument("--pairs", type=float, default=1e7, help="records to log over all seeds")
    ap.add_argument("--warmup-ms", type=float, default=100.0)
    ap.add_argument("--sim-ms", type=float, default=1000.0)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=7)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()

    seeds = seed_list(args.seeds)
    pairs = int(args.pairs)
    steps_per_run = int(round(args.sim_ms / workloads.DT))
    # ~1.05 accepted substeps per neuron-step: choose the neuron stride so
    # the runs together fill the log
    stride = max(1, math.ceil(args.neurons * 1.05 * steps_per_run * len(seeds) / pairs))
    ngpu.StartPredictorLog(pairs, stride)
    build = workloads.WORKLOADS[args.workload]
    t_all = time.perf_counter()
    for s in seeds:
        ngpu.ResetKernel()
        ngpu.SetVerbosityLevel(0)
        ngpu.SetStepPredictorMode("plain")
        t0 = time.perf_counter()
        build(ngpu, args.neurons, s)
        log("seed %d: built in %.1f s" % (s, time.perf_counter() - t0))
        ngpu.Simulate(args.warmup_ms)
        log("seed %d: calibrated and warmed up at %.1f s" % (s, time.perf_counter() - t0))
        ngpu.SetStepPredictorMode("log")
        ngpu.Simulate(args.sim_ms)
        ngpu.SetStepPredictorMode("plain")
        log("seed %d: %d records logged (stride %d) in %.1f s"
            % (s, ngpu.GetPredictorLogCount(), stride, time.perf_counter() - t0))
    ngpu.ResetKernel()
    n_log = ngpu.GetPredictorLogCount()
    sample = ngpu.GetPredictorLog(0, min(n_log, 1_000_000))
    log("log: %d records, censored fraction %.3f, target mean %.3f sd %.3f"
        % (n_log, sample[:, 9].mean(), sample[:, 8].mean(), sample[:, 8].std()))
    t0 = time.perf_counter()
    ngpu.InitPredictor(args.init_seed)
    loss = ngpu.TrainPredictor(args.steps, args.batch, args.lr, args.init_seed)
    log("trained %d steps x %d in %.1f s: loss %.4f -> %.5f"
        % (args.steps, args.batch, time.perf_counter() - t0, loss[:20].mean(), loss[-100:].mean()))
    params = ngpu.GetPredictorParams()
    out = {
        "workload": args.workload,
        "neurons": args.neurons,
        "seeds": seeds,
        "records": int(n_log),
        "stride": stride,
        "sim_ms": args.sim_ms,
        "steps": args.steps,
        "batch": args.batch,
        "lr": args.lr,
        "loss_first": float(loss[:20].mean()),
        "loss_last": float(loss[-100:].mean()),
        "wall_s": time.perf_counter() - t_all,
        "params": [float(v) for v in params],
    }
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f)
    log("saved", args.out)


if __name__ == "__main__":
    main()
