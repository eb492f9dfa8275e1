"""Synthetic benchmark networks, built through the nestgpu API as a user
script would (seeded; no data files).

scale_adex (the 1M-neuron plain-RK5 benchmark): aeif_cond_beta_multisynapse
  with 2 receptor ports (E_rev 0/-85 mV, tau_rise = tau_decay = 1 ms), 80/20
  exc/inh, fixed indegree 1000 (800 exc -> port 0, 200 inh -> port 1),
  weights U(0.025,0.075) nS exc / U(0.175,0.525) nS inh, delays U{1..50} x
  0.1 ms, one shared Poisson generator (20 kHz, 0.237 nS) tuned to ~5 Hz.

predictor_adex (the learned step-size benchmark): the same neuron with 4
  ports (fast exc/inh: tau 1/1 ms; slow exc/inh: tau_rise 2, tau_decay 10 ms;
  E_rev 0/-85/0/-85 mV), indegree 1000 (600 exc -> port 0, 200 exc -> port
  2, 150 inh -> port 1, 50 inh -> port 3; slow-port weights x0.2), and one
  Poisson generator per neuron with rate U(5,40) kHz at 0.2 nS on port 0
  (mean ~9 Hz, a mix of silent and fast-firing neurons).
"""
import ctypes

import numpy as np

DT = 0.1


def _connect_indegree(ngpu, src, tgt, k, port, w_lo, w_hi, rng):
    m = tgt.n * k
    w = rng.random(m, dtype=np.float32)
    w *= np.float32(w_hi - w_lo)
    w += np.float32(w_lo)
    d = rng.integers(1, 51, size=m, dtype=np.int16).astype(np.float32)
    d *= np.float32(DT)
    lib = ngpu._nestgpu
    ngpu.ConnSpecInit()
    ngpu.SynSpecInit()
    ngpu.SetConnSpecParam("rule", 3)           # fixed_indegree
    ngpu.SetConnSpecParam("indegree", k)
    ngpu.SetSynSpecIntParam("receptor", port)
    lib.NESTGPU_SetSynSpecFloatPtParam(b"weight_array", w.ctypes.data_as(ctypes.c_void_p))
    lib.NESTGPU_SetSynSpecFloatPtParam(b"delay_array", d.ctypes.data_as(ctypes.c_void_p))
    lib.NESTGPU_ConnectSeqSeq(src.i0, src.n, tgt.i0, tgt.n)
    ngpu._check()


SCALE_RECEPTORS = {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0], "tau_decay": [1.0, 1.0]}
SCALE_POISSON = (20000.0, 0.237)


def scale_adex(ngpu, n=1_000_000, seed=12345):
    """Build the scale_adex network; returns the neuron NodeSeq."""
    rng = np.random.default_rng(seed)
    ngpu.SetRandomSeed(seed)
    ne = int(0.8 * n)
    neurons = ngpu.Create("aeif_cond_beta_multisynapse", n, 2)
    ngpu.SetStatus(neurons, SCALE_RECEPTORS)
    pg = ngpu.Create("poisson_generator")
    ngpu.SetStatus(pg, "rate", SCALE_POISSON[0])
    ngpu.Connect(pg, neurons, {"rule": "all_to_all"},
                 {"weight": SCALE_POISSON[1], "delay": DT, "receptor": 0})
    _connect_indegree(ngpu, neurons[0:ne], neurons, 800, 0, 0.025, 0.075, rng)
    _connect_indegree(ngpu, neurons[ne:n], neurons, 200, 1, 0.175, 0.525, rng)
    return neurons


PREDICTOR_RECEPTORS = {"E_rev": [0.0, -85.0, 0.0, -85.0], "tau_rise": [1.0, 1.0, 2.0, 2.0],
                       "tau_decay": [1.0, 1.0, 10.0, 10.0]}
PREDICTOR_POISSON_W = 0.2
PREDICTOR_RATES = (5000.0, 40000.0)


def predictor_adex(ngpu, n=1_000_000, seed=200):
    """Build the predictor_adex network; returns the neuron NodeSeq."""
    rng = np.random.default_rng(seed)
    ngpu.SetRandomSeed(seed)
    ne = int(0.8 * n)
    neurons = ngpu.Create("aeif_cond_beta_multisynapse", n, 4)
    ngpu.SetStatus(neurons, PREDICTOR_RECEPTORS)
    pg = ngpu.Create("poisson_generator", n)
    rates = rng.uniform(PREDICTOR_RATES[0], PREDICTOR_RATES[1], n).astype(np.float32)
    lib = ngpu._nestgpu
    f = lib.NESTGPU_SetNeuronScalParam
    name = b"rate"
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_float]
    i0 = pg.i0
    for i, r in enumerate(rates.tolist()):
        f(i0 + i, 1, name, r)
    ngpu._check()
    ngpu.Connect(pg, neurons, {"rule": "one_to_one"},
                 {"weight": PREDICTOR_POISSON_W, "delay": DT, "receptor": 0})
    _connect_indegree(ngpu, neurons[0:ne], neurons, 600, 0, 0.025, 0.075, rng)
    _connect_indegree(ngpu, neurons[0:ne], neurons, 200, 2, 0.005, 0.015, rng)
    _connect_indegree(ngpu, neurons[ne:n], neurons, 150, 1, 0.175, 0.525, rng)
    _connect_indegree(ngpu, neurons[ne:n], neurons, 50, 3, 0.035, 0.105, rng)
    return neurons


WORKLOADS = {"scale_adex": scale_adex, "predictor_adex": predictor_adex}
