"""The open-loop input replay (tests/replay.py) reproduces a network's
neurons exactly (no GPU): a 400-neuron AdEx network with recurrent
connections and Poisson drive runs in the oracle, then 50 of its neurons are
replayed from the full run's raster; their spike times must be identical.
This pins the replay construction (the one-step shift of the generators,
stream ids, port/weight/delay export) that the full-size GPU tests rely on.
"""
import numpy as np

import netbuild
import replay
from oracle.oracle import Oracle

SEED = 777
PW = 0.35   # Poisson weight (nS): a few Hz per neuron with the recurrent input


def test_replay_reproduces_subset():
    rng = np.random.default_rng(5)
    n = 400
    net = netbuild.c1_network(seed=11, n=n, k=40, t_rate=20000.0)
    status = netbuild.C1_RECEPTORS
    o = Oracle(seed=SEED, prec=0)
    n0 = o.create("aeif_cond_beta_multisynapse", n, 2)
    for k, v in status.items():
        for p, vv in enumerate(v):
            o.set(n0, n, k, [vv] * n, port=p)
    pg = o.create("poisson_generator", 1, 0)
    o.set(pg, 1, "rate", [net["rate"]])
    streams = 5000 + np.arange(n, dtype=np.int64)
    o.connect(np.full(n, pg, np.int32), np.arange(n, dtype=np.int32) + n0, 0,
              PW, 0.1, stream=streams)
    se, te, we, de = net["exc"]
    si, ti, wi, di = net["inh"]
    o.connect(se + n0, te + n0, 0, we, de)
    o.connect(si + n0, ti + n0, 1, wi, di)
    sim = 150.0
    o.simulate(sim)
    t_all, n_all = o.events()
    assert len(t_all) > 200

    sample = np.sort(rng.choice(n, 50, replace=False))
    pos = {int(s): k for k, s in enumerate(sample)}
    src = np.concatenate([se, si, np.full(n, pg, np.int32) - n0])
    tgt = np.concatenate([te, ti, np.arange(n, dtype=np.int32)])
    port = np.concatenate([np.zeros(len(se), np.uint8), np.ones(len(si), np.uint8),
                           np.zeros(n, np.uint8)])
    w = np.concatenate([we, wi, np.full(n, PW, np.float32)])
    d = np.concatenate([de, di, np.full(n, 0.1, np.float32)])
    st = np.concatenate([np.full(len(se) + len(si), -1, np.int64), streams])
    m = np.isin(tgt, sample)
    inc = {"src": src[m].astype(np.int64) + n0, "tgt": np.array([pos[int(x)] for x in tgt[m]], np.int32),
           "port": port[m], "w": w[m], "delay": d[m], "stream": st[m]}
    spikes = {}
    for tt, nn in zip(t_all, n_all):
        spikes.setdefault(int(nn), []).append(int(tt))
    t_r, l_r = replay.replay("aeif_cond_beta_multisynapse", 2, len(sample), status, inc, spikes,
                             {int(pg): {"rate": net["rate"]}}, sim, SEED, prec=0)
    sel = np.isin(n_all, sample + n0)
    want_t, want_n = t_all[sel], n_all[sel] - n0
    got_n = sample[l_r]
    o1 = np.lexsort((want_n, want_t))
    o2 = np.lexsort((got_n, t_r))
    assert len(want_t) > 20
    np.testing.assert_array_equal(want_t[o1], t_r[o2])
    np.testing.assert_array_equal(want_n[o1], got_n[o2])
