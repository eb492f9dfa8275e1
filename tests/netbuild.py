"""Seeded random networks built identically in this build (through the
nestgpu API) and in the CPU oracle, for parity tests and the benchmark."""
import numpy as np


def random_edges(rng, n_src, n_tgt, k, src0, tgt0, w_lo, w_hi, d_lo, d_hi, dt=0.1):
    """Fixed indegree k per target, distinct sources, uniform weights and
    integer delays U{d_lo..d_hi} * dt."""
    src = np.empty(n_tgt * k, np.int32)
    for t in range(n_tgt):
        src[t * k:(t + 1) * k] = rng.choice(n_src, size=k, replace=False)
    tgt = np.repeat(np.arange(n_tgt, dtype=np.int32), k)
    w = rng.uniform(w_lo, w_hi, size=n_tgt * k).astype(np.float32)
    d = (rng.integers(d_lo, d_hi + 1, size=n_tgt * k) * dt).astype(np.float32)
    return src + src0, tgt + tgt0, w, d


def connect_edges_ngpu(ngpu, src, tgt, port, w, d):
    ngpu.Connect([int(x) for x in src], [int(x) for x in tgt], {"rule": "one_to_one"},
                 {"weight_array": [float(x) for x in w], "delay_array": [float(x) for x in d],
                  "receptor": int(port)})


def connect_edges_fast(ngpu, src, tgt, port, w, d):
    """Same as connect_edges_ngpu through the C ABI with numpy buffers."""
    import ctypes
    lib = ngpu._nestgpu
    ngpu.ConnSpecInit()
    ngpu.SynSpecInit()
    ngpu.SetConnSpecParam("rule", 0)
    ngpu.SetSynSpecIntParam("receptor", int(port))
    w = np.ascontiguousarray(w, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    s = np.ascontiguousarray(src, np.int32)
    t = np.ascontiguousarray(tgt, np.int32)
    lib.NESTGPU_SetSynSpecFloatPtParam(b"weight_array", w.ctypes.data_as(ctypes.c_void_p))
    lib.NESTGPU_SetSynSpecFloatPtParam(b"delay_array", d.ctypes.data_as(ctypes.c_void_p))
    lib.NESTGPU_ConnectGroupGroup(s.ctypes.data_as(ctypes.c_void_p), len(s),
                                  t.ctypes.data_as(ctypes.c_void_p), len(t))
    ngpu._check()


# Receptor settings and Poisson weight of the config-1 network: E_rev 0/-85 mV
# and 1 ms beta time constants as in python/test/test_brunel_list.py:33-37;
# the Poisson weight (not fixed by the config) gives ~12 Hz at 20 kHz drive.
C1_RECEPTORS = {"E_rev": [0.0, -85.0], "tau_decay": [1.0, 1.0], "tau_rise": [1.0, 1.0]}
C1_POISSON_W = 0.2


def c1_network(seed=1, n=1000, k=100, t_rate=20000.0, poiss_w=None, iaf=False):
    """Config-1 shaped network description: n neurons (80% exc), fixed
    indegree k (80% from exc to port 0, 20% from inh to port 1), weights
    U(0.5,1.5) exc / U(2,6) inh (positive weight on the inhibitory reversal
    port), delays U{1..20}*0.1 ms, one Poisson generator per neuron."""
    rng = np.random.default_rng(seed)
    ne = int(0.8 * n)
    ke = int(0.8 * k)
    ki = k - ke
    se, te, we, de = random_edges(rng, ne, n, ke, 0, 0, 0.5, 1.5, 1, 20)
    si, ti, wi, di = random_edges(rng, n - ne, n, ki, ne, 0, 2.0, 6.0, 1, 20)
    return dict(n=n, ne=ne, exc=(se, te, we, de), inh=(si, ti, wi, di), rate=t_rate,
                poiss_w=poiss_w)


# Mixed-model network in the shape of the six-population batching workload,
# scaled down: populations of every RK5 model plus iaf_psc_exp, fixed
# indegree across all populations (80% of sources excitatory -> port 0, the
# rest -> port 1), one Poisson generator per population.  Weights are per
# target model (conductance models in nS with the inhibitory port at
# E_rev = -85 mV, current models in pA).
MIXED_POPS = [
    # model, n, ports, status, exc weight range, inh weight range, Poisson rate (Hz), Poisson weight
    ("izhikevich_cond_beta", 30, 2,
     {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0], "tau_decay": [2.0, 2.0]},
     (0.002, 0.006), (0.005, 0.015), 8000.0, 0.006),
    ("izhikevich_cond_beta", 60, 2,
     {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0], "tau_decay": [2.0, 2.0]},
     (0.002, 0.006), (0.005, 0.015), 8000.0, 0.006),
    ("aeif_psc_exp", 100, 2, {},
     (20.0, 60.0), (60.0, 180.0), 20000.0, 170.0),
    ("aeif_cond_alpha_multisynapse", 200, 4,
     {"E_rev": [0.0, -85.0, 0.0, -85.0], "tau_syn": [1.0, 1.0, 5.0, 5.0]},
     (0.3, 1.0), (1.0, 3.0), 20000.0, 0.25),
    ("aeif_cond_beta_multisynapse", 300, 2,
     {"E_rev": [0.0, -85.0], "tau_rise": [1.0, 1.0], "tau_decay": [1.0, 1.0]},
     (0.5, 1.5), (2.0, 6.0), 20000.0, 0.3),
    ("iaf_psc_exp", 150, 2, {},
     (60.0, 120.0), (-240.0, -120.0), 8000.0, 22.0),
]


def mixed_network(seed=5, k=50, d_hi=20):
    """Description of the mixed network: population list with first ids,
    and per target population the (src, tgt, port, w, d) edge arrays."""
    rng = np.random.default_rng(seed)
    pops = []
    i0 = 0
    for (model, n, npt, st, we, wi, rate, pw) in MIXED_POPS:
        pops.append(dict(model=model, n=n, n_port=npt, status=st, i0=i0, we=we, wi=wi,
                         rate=rate, pw=pw))
        i0 += n
    n_tot = i0
    n_exc = int(0.8 * n_tot)     # global ids below n_exc are excitatory sources
    edges = []
    for p in pops:
        n = p["n"]
        src = np.empty(n * k, np.int32)
        for t in range(n):
            src[t * k:(t + 1) * k] = rng.choice(n_tot, size=k, replace=False)
        tgt = np.repeat(np.arange(n, dtype=np.int32), k) + p["i0"]
        exc = src < n_exc
        port = np.where(exc, 0, 1).astype(np.uint8)
        w = np.where(exc, rng.uniform(*p["we"], size=n * k),
                     rng.uniform(*p["wi"], size=n * k)).astype(np.float32)
        d = (rng.integers(1, d_hi + 1, size=n * k) * 0.1).astype(np.float32)
        edges.append((src, tgt, port, w, d))
    return dict(pops=pops, n=n_tot, edges=edges)
