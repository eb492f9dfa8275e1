"""Open-loop input replay of part of a network in the CPU oracle.

A network too large for the oracle is simulated in full by this build; for a
sample of its neurons the oracle then simulates only those neurons, fed with
exactly the input they received: every synapse into them (source, port,
weight, delay as exported by NESTGPU_AMD_GetIncoming), each spiking source
replaced by a spike_generator that re-emits the source's recorded spikes,
and each poisson_generator source kept as a generator with the same Poisson
stream ids (the synapse's index in the source-major table), so the draws
are the same.  The sampled neurons' spike times in the oracle can then be
compared with the full run's: identical for the float path, and the FP64
path states the error of the float one.

A spike recorded at time index T (time T*dt, the end of its step) is
re-emitted by a spike_generator at time (T-1)*dt: the generator emits in the
step that starts at its spike time (src/spike_generator.cu:49-67), i.e. the
step with time index T, the original emission step.
"""
import numpy as np

DT = 0.1


def replay(model, n_port, n_tgt, status, inc, spikes, poisson, sim_ms, seed, prec=0):
    """Simulate n_tgt neurons of `model` in the oracle with the replayed input.

    status: name -> scalar, per-target array, or (for port parameters) list
            of per-port values
    inc:    dict of arrays src (global id), tgt (0..n_tgt-1), port, w, delay,
            stream (Poisson stream id; used for poisson sources)
    spikes: global source id -> sorted array of emission time indices
    poisson: global source id -> dict(rate, origin, start, stop)
    Returns (t_idx, local target) of the targets' spikes, sorted.
    """
    from oracle.oracle import Oracle
    o = Oracle(seed=seed, prec=prec)
    n0 = o.create(model, n_tgt, n_port)
    for k, v in status.items():
        if isinstance(v, list):
            for p, vv in enumerate(v):
                o.set(n0, n_tgt, k, np.broadcast_to(np.float32(vv), (n_tgt,)), port=p)
        else:
            o.set(n0, n_tgt, k, np.broadcast_to(np.asarray(v, np.float32), (n_tgt,)))
    src = np.asarray(inc["src"])
    is_pg = np.isin(src, np.fromiter(poisson.keys(), np.int64, len(poisson)))
    # spiking sources that reach the sample and spiked at all
    sp_src = np.unique(src[~is_pg])
    sp_src = np.array([s for s in sp_src if s in spikes and len(spikes[s]) > 0], np.int64)
    local = {}
    if len(sp_src):
        g0 = o.create("spike_generator", len(sp_src), 0)
        for k, s in enumerate(sp_src):
            t = (np.asarray(spikes[s], np.int64) - 1).astype(np.float64) * DT
            o.set_spike_times(g0 + k, np.asarray(t, np.float32))
            local[int(s)] = g0 + k
    pg_src = np.unique(src[is_pg])
    for s in pg_src:
        p = poisson[int(s)]
        g = o.create("poisson_generator", 1, 0)
        for key in ("rate", "origin", "start", "stop"):
            if key in p:
                o.set(g, 1, key, [p[key]])
        local[int(s)] = g
    keep = np.array([int(s) in local for s in src], bool)
    lsrc = np.array([local[int(s)] for s in src[keep]], np.int32)
    o.connect(lsrc, np.asarray(inc["tgt"])[keep] + n0, np.asarray(inc["port"])[keep],
              np.asarray(inc["w"])[keep], np.asarray(inc["delay"])[keep],
              stream=np.where(is_pg[keep], np.asarray(inc["stream"])[keep], -1))
    o.simulate(sim_ms)
    t, nd = o.events()
    m = (nd >= n0) & (nd < n0 + n_tgt)
    return t[m], nd[m] - n0


def spike_steps(times_ms):
    """Recorded spike times (ms, the end of the emission step) -> time indices."""
    return np.rint(np.asarray(times_ms, np.float64) / DT).astype(np.int64)
