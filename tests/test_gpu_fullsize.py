"""The benchmark networks at full size on the MI355X.

  C3 scale_adex: 1,000,000 aeif_cond_beta_multisynapse, 2 ports, indegree 1000
     (1e9 synapses), delays U{1..50} x 0.1 ms, shared Poisson drive
  C5 predictor_adex: the same with 4 ports and one Poisson generator per
     neuron at U(5,40) kHz
  C2 microcircuit: 77,169 iaf_psc_exp in 8 populations, ~4.5e8 synapses

Each runs in full on the GPU and is checked by properties that hold at any
size, and by the oracle on a sample of its neurons:
  * no spike was dropped from a spike list;
  * every neuron's spike_count equals the number of its recorded spike times
    (python/test/test_spike_times.py);
  * the delivered (spike, synapse) events equal sum_i spikes_i x out-degree_i
    (every spike reaches each of its synapses exactly once);
  * the population rate lies in the stated band;
  * 1000 sampled neurons replayed in the oracle's float path with exactly
    the input they received (tests/replay.py) spike at identical time steps;
    for C3 the FP64 oracle replays them too, stating the float path's error.
"""
import ctypes
import json
import os
import time

import numpy as np
import pytest

import replay
import workloads
from conftest import gpu_available

pytestmark = pytest.mark.gpu

DT = 0.1


@pytest.fixture(scope="module")
def ngpu():
    if not gpu_available():
        pytest.skip("no GPU")
    import nestgpu
    nestgpu.SetVerbosityLevel(0)
    return nestgpu


@pytest.fixture(autouse=True)
def fresh(ngpu):
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    yield
    ngpu.ResetKernel()


def log(*a):
    print(*a, flush=True)


def _out_degrees(ngpu, i0, n):
    lib = ngpu._nestgpu
    f = lib.NESTGPU_AMD_GetOutDegrees
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    out = np.empty(n, np.int64)
    f(i0, n, out.ctypes.data)
    ngpu._check()
    return out


def _incoming(ngpu, targets):
    lib = ngpu._nestgpu
    t = np.ascontiguousarray(targets, np.int32)
    cnt = lib.NESTGPU_AMD_CountIncoming
    cnt.argtypes = [ctypes.c_void_p, ctypes.c_int]
    cnt.restype = ctypes.c_longlong
    m = cnt(t.ctypes.data, len(t))
    ngpu._check()
    a = {"src": np.empty(m, np.int32), "tgt": np.empty(m, np.int32), "port": np.empty(m, np.uint8),
         "w": np.empty(m, np.float32), "delay": np.empty(m, np.float32),
         "stream": np.empty(m, np.int64)}
    g = lib.NESTGPU_AMD_GetIncoming
    g.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 6
    g(t.ctypes.data, len(t), *(a[k].ctypes.data for k in ("src", "tgt", "port", "w", "delay", "stream")))
    ngpu._check()
    return a


def _spike_steps(ngpu, pops):
    return [replay.spike_steps(ts) for pop in pops for ts in ngpu.GetRecSpikeTimes(pop)]


def _pops(ngpu, neurons, sizes=None):
    """The node groups a run spans (per-node getters work group by group)."""
    if sizes is None:
        return [neurons]
    out, i0 = [], neurons.i0
    for m in sizes:
        out.append(ngpu.NodeSeq(i0, m))
        i0 += m
    return out


def _check_full_run(ngpu, neurons, sim_ms, max_rec, rate_band, sizes=None):
    """Runs sim_ms and checks the size-independent properties; returns the
    per-neuron spike steps."""
    n = neurons.n
    pops = _pops(ngpu, neurons, sizes)
    st0 = ngpu.GetStepStats()
    t0 = time.perf_counter()
    ngpu.Simulate(sim_ms)
    ngpu.Synchronize()
    log("simulated %.0f ms in %.2f s" % (sim_ms, time.perf_counter() - t0))
    st1 = ngpu.GetStepStats()
    assert st1["dropped_spikes"] == 0
    steps = _spike_steps(ngpu, pops)
    counts = np.concatenate([np.asarray(ngpu.GetStatus(pop, "spike_count"), np.int64).reshape(-1)
                             for pop in pops])
    n_rec = np.array([len(s) for s in steps], np.int64)
    full = n_rec < max_rec - 1          # recording not saturated
    assert full.mean() > 0.999
    np.testing.assert_array_equal(counts[full], n_rec[full])
    outdeg = _out_degrees(ngpu, neurons.i0, n)
    events = st1["syn_events"] - st0["syn_events"]
    assert events == int(np.dot(counts, outdeg)), (events, int(np.dot(counts, outdeg)))
    assert st1["spikes"] - st0["spikes"] == counts.sum()
    rate = counts.sum() / n / (sim_ms * 1e-3)
    log("rate %.2f Hz, %d spikes, %d synaptic events" % (rate, counts.sum(), events))
    assert rate_band[0] < rate < rate_band[1]
    return steps, counts


def _replay_sample(ngpu, model, n_port, neurons, steps, sample, status, poisson, sim_ms, seed,
                   prec=0):
    inc = _incoming(ngpu, neurons.i0 + sample)
    pos = np.full(neurons.i0 + neurons.n, -1, np.int64)
    pos[neurons.i0 + sample] = np.arange(len(sample))
    inc["tgt"] = pos[inc["tgt"]].astype(np.int32)
    spikes = {neurons.i0 + i: s for i, s in enumerate(steps) if len(s)}
    t0 = time.perf_counter()
    t, l = replay.replay(model, n_port, len(sample), status, inc, spikes, poisson, sim_ms, seed,
                         prec=prec)
    log("replayed %d neurons (%d synapses in) in the oracle (prec %d) in %.1f s"
        % (len(sample), len(inc["src"]), prec, time.perf_counter() - t0))
    return t, l


def _gpu_sample_events(steps, sample):
    t = np.concatenate([steps[i] for i in sample] + [np.zeros(0, np.int64)])
    l = np.concatenate([np.full(len(steps[i]), k, np.int64) for k, i in enumerate(sample)]
                       + [np.zeros(0, np.int64)])
    o = np.lexsort((l, t))
    return t[o], l[o]


def _sorted(t, l):
    o = np.lexsort((l, t))
    return np.asarray(t)[o], np.asarray(l)[o]


def _match_fraction(ta, la, tb, lb):
    a = set(zip(ta.tolist(), la.tolist()))
    b = set(zip(tb.tolist(), lb.tolist()))
    return len(a & b) / max(len(a | b), 1)


def test_c3_scale_adex_full(ngpu):
    n, seed, sim_ms, max_rec = 1_000_000, 12345, 200.0, 100
    t0 = time.perf_counter()
    neurons = workloads.scale_adex(ngpu, n, seed)
    ngpu.ActivateSpikeCount(neurons)
    ngpu.ActivateRecSpikeTimes(neurons, max_rec)
    ngpu.Calibrate()
    log("C3 built and calibrated in %.1f s" % (time.perf_counter() - t0))
    steps, counts = _check_full_run(ngpu, neurons, sim_ms, max_rec, (1.0, 40.0))
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(n, 1000, replace=False))
    pg = neurons.i0 + n        # the shared generator, created after the neurons
    poisson = {pg: {"rate": workloads.SCALE_POISSON[0]}}
    status = dict(workloads.SCALE_RECEPTORS)
    want = _gpu_sample_events(steps, sample)
    assert len(want[0]) > 100
    got = _sorted(*_replay_sample(ngpu, "aeif_cond_beta_multisynapse", 2, neurons, steps, sample,
                                  status, poisson, sim_ms, seed + 12345, prec=0))
    np.testing.assert_array_equal(want[0], got[0])
    np.testing.assert_array_equal(want[1], got[1])
    # the FP64 integrator on the same input: the float path's error
    g64 = _sorted(*_replay_sample(ngpu, "aeif_cond_beta_multisynapse", 2, neurons, steps, sample,
                                  status, poisson, sim_ms, seed + 12345, prec=1))
    frac = _match_fraction(want[0], want[1], g64[0], g64[1])
    log("C3 sample: %d spikes; FP32 vs FP64 oracle: %.4f of spikes at the same step, "
        "%d vs %d spikes" % (len(want[0]), frac, len(want[0]), len(g64[0])))
    assert frac > 0.98


def test_c5_predictor_adex_full(ngpu):
    n, seed, sim_ms, max_rec = 1_000_000, 200, 200.0, 100
    t0 = time.perf_counter()
    neurons = workloads.predictor_adex(ngpu, n, seed)
    ngpu.ActivateSpikeCount(neurons)
    ngpu.ActivateRecSpikeTimes(neurons, max_rec)
    ngpu.Calibrate()
    log("C5 built and calibrated in %.1f s" % (time.perf_counter() - t0))
    # per-neuron generator rates (one generator per neuron, created after them)
    pg0 = neurons.i0 + n
    rates = np.asarray(ngpu.GetStatus(ngpu.NodeSeq(pg0, n), "rate"), np.float32).reshape(-1)
    steps, counts = _check_full_run(ngpu, neurons, sim_ms, max_rec, (1.0, 60.0))
    rng = np.random.default_rng(2)
    sample = np.sort(rng.choice(n, 1000, replace=False))
    poisson = {pg0 + i: {"rate": float(rates[i])} for i in sample}
    status = dict(workloads.PREDICTOR_RECEPTORS)
    want = _gpu_sample_events(steps, sample)
    assert len(want[0]) > 100
    got = _sorted(*_replay_sample(ngpu, "aeif_cond_beta_multisynapse", 4, neurons, steps, sample,
                                  status, poisson, sim_ms, seed + 12345, prec=0))
    np.testing.assert_array_equal(want[0], got[0])
    np.testing.assert_array_equal(want[1], got[1])


def _window_stats(times, t0_ms, t1_ms, n):
    """Mean rate over all neurons and mean ISI CV over the neurons with at
    least 4 spikes inside [t0, t1)."""
    cnt, cvs = 0, []
    for ts in times:
        a = np.asarray(ts, np.float64)
        a = a[(a >= t0_ms) & (a < t1_ms)]
        cnt += a.size
        if a.size >= 4:
            isi = np.diff(np.sort(a))
            cvs.append(isi.std() / isi.mean())
    return cnt / n / ((t1_ms - t0_ms) * 1e-3), float(np.mean(cvs)), len(cvs)


def test_c5_learned_vs_plain_statistics(ngpu):
    # the learned step-size proposal on the full config-5 network (the
    # committed predictor, trained on logs of seeds 100-109; this is seed
    # 200): windows plain / learned / plain of 400 ms after 300 ms settling.
    # The learned window's mean rate and mean ISI CV lie within 2% of the
    # mean of the two plain windows around it (at 1e6 neurons, realisation
    # noise of both statistics is far below that), with far fewer rejected
    # attempts
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "nest-gpu_amd", "predictors", "predictor_adex.json")) as f:
        pred = json.load(f)
    n, seed, win = 1_000_000, 200, 400.0
    neurons = workloads.predictor_adex(ngpu, n, seed)
    ngpu.ActivateRecSpikeTimes(neurons, 250)
    ngpu.Calibrate()
    try:
        ngpu.SetStepPredictorMode("plain")
        ngpu.SetPredictorParams(np.asarray(pred["params"], np.float32))
        ngpu.Simulate(300.0)
        t = 300.0
        wins = []
        for mode in ("plain", "learned", "plain"):
            ngpu.SetStepPredictorMode(mode)
            s0 = ngpu.GetStepStats()
            ngpu.Simulate(win)
            s1 = ngpu.GetStepStats()
            wins.append((mode, t, t + win, {k: s1[k] - s0[k] for k in ("attempts", "rejected")}))
            t += win
    finally:
        ngpu.SetStepPredictorMode("plain")
    t0 = time.perf_counter()
    times = ngpu.GetRecSpikeTimes(neurons)
    stats = [_window_stats(times, a, b, n) for (_, a, b, _) in wins]
    log("spike statistics in %.1f s" % (time.perf_counter() - t0))
    for (mode, a, b, st), (r, cv, ncv) in zip(wins, stats):
        log("%-7s %5.0f-%5.0f ms: rate %.4f Hz, ISI CV %.5f over %d neurons, rejected %.5f"
            % (mode, a, b, r, cv, ncv, st["rejected"] / st["attempts"]))
    r_ref = 0.5 * (stats[0][0] + stats[2][0])
    cv_ref = 0.5 * (stats[0][1] + stats[2][1])
    assert abs(stats[1][0] - r_ref) <= 0.02 * r_ref
    assert abs(stats[1][1] - cv_ref) <= 0.02 * cv_ref
    rej_plain = 0.5 * sum(w[3]["rejected"] / w[3]["attempts"] for w in (wins[0], wins[2]))
    assert wins[1][3]["rejected"] / wins[1][3]["attempts"] < 0.5 * rej_plain


def test_c2_microcircuit_full(ngpu):
    seed, sim_ms, max_rec = 7, 100.0, 400
    t0 = time.perf_counter()
    neurons = workloads.microcircuit(ngpu, None, seed)
    n = neurons.n
    v0 = np.concatenate([np.asarray(ngpu.GetStatus(pop, "V_m_rel"), np.float32).reshape(-1)
                         for pop in _pops(ngpu, neurons, workloads.MC_SIZES)])
    for pop in _pops(ngpu, neurons, workloads.MC_SIZES):
        ngpu.ActivateSpikeCount(pop)
        ngpu.ActivateRecSpikeTimes(pop, max_rec)
    ngpu.Calibrate()
    log("C2 built (%d synapses) and calibrated in %.1f s"
        % (workloads.microcircuit.n_syn, time.perf_counter() - t0))
    steps, counts = _check_full_run(ngpu, neurons, sim_ms, max_rec, (1.0, 500.0),
                                    sizes=workloads.MC_SIZES)
    rng = np.random.default_rng(3)
    sample = np.sort(rng.choice(n, 1000, replace=False))
    # the 8 generators follow the neurons, one per population in order
    pg0 = neurons.i0 + n
    poisson = {pg0 + k: {"rate": 8.0 * kx} for k, kx in enumerate(workloads.MC_K_EXT)}
    status = dict(workloads.MC_PARAMS)
    status["V_m_rel"] = v0[sample]
    want = _gpu_sample_events(steps, sample)
    assert len(want[0]) > 100
    got = _sorted(*_replay_sample(ngpu, "iaf_psc_exp", 2, neurons, steps, sample,
                                  status, poisson, sim_ms, seed + 12345, prec=0))
    np.testing.assert_array_equal(want[0], got[0])
    np.testing.assert_array_equal(want[1], got[1])
