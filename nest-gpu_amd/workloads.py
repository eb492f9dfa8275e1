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

batching (six populations of five models in one launch per step): 300 and
  3,000 izhikevich_cond_beta (2 ports), 30,000 aeif_psc_exp, 120,000
  aeif_cond_alpha_multisynapse (4 ports), 300,000 aeif_cond_beta_multisynapse
  (2 ports) and 50,000 iaf_psc_exp (503,300 neurons); the first 80% of the
  ids are the excitatory sources; every neuron has fixed indegree 500 (400
  excitatory sources -> port 0, 100 inhibitory -> port 1) over all
  populations, 2.5e8 synapses; weights normal with sd 10% of the mean, the
  inhibitory mean 4x the excitatory one (as the microcircuit), the means per
  target model (nS for conductance ports, pA for current ports; inhibitory
  current weights negative for iaf_psc_exp, positive on aeif_psc_exp's
  inhibitory port, which subtracts); delays N(1.5, 0.75) ms exc, N(0.75,
  0.375) ms inh, at least 0.1 ms; one Poisson generator per population.

microcircuit (the delivery-bound iaf_psc_exp benchmark): 8 populations of
  the cortical microcircuit's sizes (77,169 neurons; microcircuit neuron
  parameters of python/Potjans_2014_s/network_params.py: E_L -65 mV, V_th
  -50 mV, V_reset -65 mV, C_m 250 pF, tau_m 10 ms, tau_syn 0.5 ms, t_ref
  2 ms); each population pair connected by fixed_total_number with
  K = round(p N_pre N_post), p ~ U(0.01, 0.15) per pair (64 seeded draws);
  weights N(87.8, 8.78) pA from excitatory populations, x -4 from
  inhibitory ones (kept on their sign); delays N(1.5, 0.75) ms exc,
  N(0.75, 0.375) ms inh, at least 0.1 ms; Poisson input 8 Hz x K_ext per
  neuron (K_ext = 1600, 1500, 2100, 1900, 2000, 1900, 2900, 2100) at 87.8 pA.
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


# model, size, ports, status, excitatory mean weight, inhibitory sign,
# Poisson rate (Hz), Poisson weight
BATCH_POPS = [
    ("izhikevich_cond_beta", 300, 2,
     {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0], "tau_decay": [2.0, 2.0]}, 0.0004, 1.0,
     8000.0, 0.006),
    ("izhikevich_cond_beta", 3000, 2,
     {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0], "tau_decay": [2.0, 2.0]}, 0.0004, 1.0,
     8000.0, 0.006),
    ("aeif_psc_exp", 30000, 2, {}, 4.0, 1.0, 20000.0, 170.0),
    ("aeif_cond_alpha_multisynapse", 120000, 4,
     {"E_rev": [0.0, -85.0, 0.0, -85.0], "tau_syn": [1.0, 1.0, 5.0, 5.0]}, 0.065, 1.0,
     20000.0, 0.25),
    ("aeif_cond_beta_multisynapse", 300000, 2,
     {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0], "tau_decay": [1.0, 1.0]}, 0.1, 1.0,
     20000.0, 0.3),
    ("iaf_psc_exp", 50000, 2, {}, 9.0, -1.0, 8000.0, 22.0),
]
BATCH_K = (400, 100)


def batching(ngpu, n=None, seed=300):
    """Build the six-population batching network (n is ignored: the
    population sizes are fixed); returns the NodeSeq of all its neurons."""
    rng = np.random.default_rng(seed)
    ngpu.SetRandomSeed(seed)
    pops = []
    for (model, m, npt, st, _, _, _, _) in BATCH_POPS:
        pop = ngpu.Create(model, m, npt)
        if st:
            ngpu.SetStatus(pop, st)
        pops.append(pop)
    i0 = pops[0].i0
    n_tot = sum(p[1] for p in BATCH_POPS)
    n_exc = int(0.8 * n_tot)
    exc = ngpu.NodeSeq(i0, n_exc)
    inh = ngpu.NodeSeq(i0 + n_exc, n_tot - n_exc)
    lib = ngpu._nestgpu
    for pop, (model, m, npt, st, we, sgn, rate, pw) in zip(pops, BATCH_POPS):
        for src, k, port, mean, d_mu, d_sd in ((exc, BATCH_K[0], 0, we, 1.5, 0.75),
                                               (inh, BATCH_K[1], 1, 4.0 * we * sgn, 0.75, 0.375)):
            cnt = m * k
            w = rng.normal(mean, abs(mean) * 0.1, cnt).astype(np.float32)
            w = np.maximum(w, 0.0) if mean > 0 else np.minimum(w, 0.0)
            d = np.maximum(rng.normal(d_mu, d_sd, cnt), DT).astype(np.float32)
            ngpu.ConnSpecInit()
            ngpu.SynSpecInit()
            ngpu.SetConnSpecParam("rule", 3)           # fixed_indegree
            ngpu.SetConnSpecParam("indegree", k)
            ngpu.SetSynSpecIntParam("receptor", port)
            lib.NESTGPU_SetSynSpecFloatPtParam(b"weight_array", w.ctypes.data_as(ctypes.c_void_p))
            lib.NESTGPU_SetSynSpecFloatPtParam(b"delay_array", d.ctypes.data_as(ctypes.c_void_p))
            lib.NESTGPU_ConnectSeqSeq(src.i0, src.n, pop.i0, pop.n)
            ngpu._check()
    for pop, (model, m, npt, st, we, sgn, rate, pw) in zip(pops, BATCH_POPS):
        pg = ngpu.Create("poisson_generator")
        ngpu.SetStatus(pg, "rate", rate)
        ngpu.Connect(pg, pop, {"rule": "all_to_all"}, {"weight": pw, "delay": DT, "receptor": 0})
    batching.pops = pops
    batching.n_syn = n_tot * sum(BATCH_K)
    return ngpu.NodeSeq(i0, n_tot)


MC_SIZES = [20683, 5834, 21915, 5479, 4850, 1065, 14395, 2948]
MC_K_EXT = [1600, 1500, 2100, 1900, 2000, 1900, 2900, 2100]
MC_PARAMS = {"E_L": -65.0, "Theta_rel": 15.0, "V_reset_rel": 0.0, "C_m": 250.0, "tau_m": 10.0,
             "tau_ex": 0.5, "tau_in": 0.5, "t_ref": 2.0}


def microcircuit(ngpu, n=None, seed=7):
    """Build the microcircuit-shaped network (n is ignored: the population
    sizes are fixed); returns the NodeSeq of all its neurons."""
    rng = np.random.default_rng(seed)
    ngpu.SetRandomSeed(seed)
    pops = [ngpu.Create("iaf_psc_exp", m, 2) for m in MC_SIZES]
    for pop in pops:
        ngpu.SetStatus(pop, MC_PARAMS)
        v0 = rng.uniform(0.0, 15.0, pop.n).astype(np.float32)
        ngpu.SetStatus(pop, "V_m_rel", {"array": [float(v) for v in v0]})
    prob = rng.uniform(0.01, 0.15, (8, 8))
    lib = ngpu._nestgpu
    n_syn = 0
    for j, pre in enumerate(pops):            # presynaptic population
        exc = j % 2 == 0
        for i, post in enumerate(pops):
            k = int(round(prob[i, j] * pre.n * post.n))
            if k == 0:
                continue
            if exc:
                w = np.maximum(rng.normal(87.8, 8.78, k), 0.0).astype(np.float32)
                d = np.maximum(rng.normal(1.5, 0.75, k), DT).astype(np.float32)
            else:
                w = np.minimum(rng.normal(-4 * 87.8, 4 * 8.78, k), 0.0).astype(np.float32)
                d = np.maximum(rng.normal(0.75, 0.375, k), DT).astype(np.float32)
            ngpu.ConnSpecInit()
            ngpu.SynSpecInit()
            ngpu.SetConnSpecParam("rule", 2)       # fixed_total_number
            ngpu.SetConnSpecParam("total_num", k)
            ngpu.SetSynSpecIntParam("receptor", 0 if exc else 1)
            lib.NESTGPU_SetSynSpecFloatPtParam(b"weight_array", w.ctypes.data_as(ctypes.c_void_p))
            lib.NESTGPU_SetSynSpecFloatPtParam(b"delay_array", d.ctypes.data_as(ctypes.c_void_p))
            lib.NESTGPU_ConnectSeqSeq(pre.i0, pre.n, post.i0, post.n)
            ngpu._check()
            n_syn += k
    for pop, kx in zip(pops, MC_K_EXT):
        pg = ngpu.Create("poisson_generator")
        ngpu.SetStatus(pg, "rate", 8.0 * kx)
        ngpu.Connect(pg, pop, {"rule": "all_to_all"}, {"weight": 87.8, "delay": 1.5, "receptor": 0})
    microcircuit.n_syn = n_syn
    return ngpu.NodeSeq(pops[0].i0, sum(MC_SIZES))


WORKLOADS = {"scale_adex": scale_adex, "predictor_adex": predictor_adex, "microcircuit": microcircuit,
             "batching": batching}
