"""Parity of the HIP simulation step with the CPU oracle (MI355X only).

All cases run in this one process; nestgpu.ResetKernel() separates the
networks.  Tolerances: integer/routing work bit-exact; single-neuron RK5
trajectories within 1e-5 relative of the oracle's float path; iaf_psc_exp and
Poisson input bit-exact (same float operation sequence, contraction off).
"""
import math
import os
import sys

import numpy as np
import pytest

from conftest import gpu_available
from scenarios import SCENARIOS, NestGpuBackend, OracleBackend, nest_trace, trace_rmse
import netbuild

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ngpu():
    if not gpu_available():
        pytest.skip("no GPU")
    import nestgpu
    nestgpu.SetVerbosityLevel(0)
    return nestgpu


@pytest.fixture(autouse=True)
def fresh(ngpu):
    print("case start", flush=True)
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    yield


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_single_neuron_trace(ngpu, name):
    sc = SCENARIOS[name]
    t, v = NestGpuBackend(ngpu).run_single(sc)
    _, vo = OracleBackend(0).run_single(sc)
    _, v64 = OracleBackend(1).run_single(sc)
    _, vn = nest_trace(sc["golden"])
    assert len(v) == len(vo) == 8001
    # reference acceptance (python/test/test_*.py): rmse vs NEST trace
    assert trace_rmse(v, vn) < sc["tol"]
    rel = np.max(np.abs(v - vo) / np.maximum(np.abs(vo), 1e-6))
    print(name, "max rel diff vs oracle float", rel,
          "vs FP64", np.max(np.abs(v - v64) / np.abs(v64)))
    # within 1e-5 relative of the reference float path, and in fact the same
    # float operation sequence as the oracle's: identical values
    assert rel < 1e-5
    np.testing.assert_array_equal(v, vo)


# (model, ports): the reference's maximum of 20 receptor ports
# (src/aeif_cond_beta_multisynapse.h:165) for the multisynapse models, and
# aeif_psc_delta, whose ports all jump V_m (applied port after port)
MANY_PORTS = [("aeif_cond_beta_multisynapse", 20), ("aeif_cond_alpha_multisynapse", 20),
              ("izhikevich_cond_beta", 20), ("aeif_psc_alpha_multisynapse", 20),
              ("aeif_psc_exp_multisynapse", 20), ("aeif_psc_delta", 3),
              ("aeif_cond_beta_multisynapse", 9)]


def _many_ports_case(model, n_port):
    n = 3
    cond = "cond" in model
    izh = model.startswith("izhikevich")
    # one generator per port; several ports' inputs arrive in the same step
    times = [[5.0 + 3.0 * (k % 7), 40.0 + 2.0 * (k % 4), 60.0 + (k % 2)] for k in range(n_port)]
    delay = [0.1 * (1 + k % 3) for k in range(n_port)]

    def weight(k, i):
        if cond:
            return (0.4 + 0.05 * k) * (1.0 + 0.5 * i) * (4.0 if izh else 1.0)
        sign = 1.0 if k % 2 == 0 else -1.0
        scale = 1.0 if model == "aeif_psc_delta" else 25.0
        return sign * scale * (1.0 + 0.1 * k) * (1.0 + 0.5 * i)

    status = {"I_e": 10.0 if izh else 700.0}
    if cond:
        status["E_rev"] = [0.0 if k % 2 == 0 else -85.0 for k in range(n_port)]
    return n, times, delay, weight, status


@pytest.mark.parametrize("model,n_port", MANY_PORTS)
def test_many_ports_bitexact(ngpu, model, n_port):
    from oracle.oracle import Oracle
    n, times, delay, weight, status = _many_ports_case(model, n_port)
    print("case", model, n_port, flush=True)
    nrn = ngpu.Create(model, n, n_port)
    ngpu.SetStatus(nrn, status)
    sg = ngpu.Create("spike_generator", n_port)
    for k in range(n_port):
        ngpu.SetStatus([sg[k]], {"spike_times": times[k]})
    for k in range(n_port):
        for i in range(n):
            ngpu.Connect([sg[k]], [nrn[i]], {"rule": "one_to_one"},
                         {"receptor": k, "weight": weight(k, i), "delay": delay[k]})
    wname = "u" if model.startswith("izhikevich") else "w"
    rec = ngpu.CreateRecord("", ["V_m"] * n + [wname] * n, [nrn[i] for i in range(n)] * 2, [0] * (2 * n))
    ngpu.Simulate(100.0)
    g = np.array(ngpu.GetRecordData(rec), np.float32)[:, 1:]

    o = Oracle(prec=0)
    n0 = o.create(model, n, n_port)
    for key, v in status.items():
        if isinstance(v, list):
            for k, vv in enumerate(v):
                o.set(n0, n, key, [vv] * n, port=k)
        else:
            o.set(n0, n, key, [v] * n)
    s0 = o.create("spike_generator", n_port, 0)
    for k in range(n_port):
        o.set_spike_times(s0 + k, times[k])
    for k in range(n_port):
        for i in range(n):
            o.connect([s0 + k], [n0 + i], [k], [weight(k, i)], [delay[k]])
    recs = [o.record(n0 + i, "V_m") for i in range(n)] + [o.record(n0 + i, wname) for i in range(n)]
    o.simulate(100.0)
    ref = np.stack([o.record_data(r)[1] for r in recs], axis=1)
    print(model, n_port, "V range", float(ref[:, :n].min()), float(ref[:, :n].max()), flush=True)
    assert g.shape == ref.shape
    # inputs changed the trajectory (every port reached the neuron)
    assert np.ptp(ref[:, 0]) > 1.0
    np.testing.assert_array_equal(g, ref)


def test_izhikevich_cond_beta_single(ngpu):
    from oracle.oracle import Oracle
    status = {"I_e": 5.0}
    res = []
    for be in ("gpu", "orc"):
        if be == "gpu":
            n = ngpu.Create("izhikevich_cond_beta", 1, 2)
            ngpu.SetStatus(n, {"I_e": 5.0, "E_rev": [0.0, -85.0]})
            sg = ngpu.Create("spike_generator")
            ngpu.SetStatus(sg, {"spike_times": [20.0, 50.0, 51.0]})
            ngpu.Connect(sg, n, {"rule": "all_to_all"}, {"receptor": 0, "weight": 2.0, "delay": 1.0})
            rec = ngpu.CreateRecord("", ["V_m", "u"], [n[0], n[0]], [0, 0])
            ngpu.Simulate(200.0)
            d = np.array(ngpu.GetRecordData(rec))
            res.append((d[:, 1], d[:, 2]))
        else:
            o = Oracle(prec=0)
            n = o.create("izhikevich_cond_beta", 1, 2)
            o.set(n, 1, "I_e", [5.0])
            o.set(n, 1, "E_rev", [0.0], 0)
            o.set(n, 1, "E_rev", [-85.0], 1)
            sg = o.create("spike_generator", 1, 0)
            o.set_spike_times(sg, [20.0, 50.0, 51.0])
            o.connect([sg], [n], [0], [2.0], [1.0])
            r1, r2 = o.record(n, "V_m"), o.record(n, "u")
            o.simulate(200.0)
            res.append((o.record_data(r1)[1], o.record_data(r2)[1]))
    (vg, ug), (vo, uo) = res
    assert np.max(np.abs(vg - vo) / np.maximum(np.abs(vo), 1.0)) < 1e-5
    assert np.max(np.abs(ug - uo) / np.maximum(np.abs(uo), 1.0)) < 1e-5


def test_spike_detector_known_answer(ngpu):
    # python/test/test_spike_detector.py:5-51
    neuron = ngpu.Create("aeif_cond_beta_multisynapse", 3)
    ngpu.SetStatus(neuron, {"I_e": 1000.0})
    sd = ngpu.Create("spike_detector")
    for k, (w, d) in enumerate([(1.0, 10.0), (2.0, 20.0), (3.0, 30.0)]):
        ngpu.Connect([neuron[k]], sd, {"rule": "one_to_one"}, {"weight": w, "delay": d, "receptor": 0})
    rn = ngpu.CreateRecord("", ["spike"], [neuron[0]], [0])
    rs = ngpu.CreateRecord("", ["spike_height"], [sd[0]], [0])
    ngpu.Simulate()
    sn = [r[1] for r in ngpu.GetRecordData(rn)]
    ss = [r[1] for r in ngpu.GetRecordData(rs)]
    hits = 0
    for i in range(len(sn) - 400):
        if sn[i] > 0.5:
            hits += 1
            assert abs(ss[i + 101] - 1.0) < 1e-6
            assert abs(ss[i + 201] - 2.0) < 1e-6
            assert abs(ss[i + 301] - 3.0) < 1e-6
    assert hits > 5


def test_routing_brute_force(ngpu):
    # every (spike, synapse): arrival one delay after emission, on its port,
    # with weight*height; multiple ports and delay groups per source
    times = [[1.0, 2.5, 7.0], [3.0, 3.1]]
    heights = [[1.0, 1.0, 1.0], [0.5, 2.0]]
    sg = ngpu.Create("spike_generator", 2)
    det = ngpu.Create("spike_detector", 6)
    for k in range(2):
        ngpu.SetStatus([sg[k]], {"spike_times": times[k], "spike_heights": heights[k]})
    conns = []
    for k in range(2):
        for j, d in enumerate([0.1, 0.4, 1.3]):
            w = 0.25 * (j + 1)
            conns.append((k, j, w, d))
            ngpu.Connect([sg[k]], [det[3 * k + j]], {"rule": "one_to_one"}, {"weight": w, "delay": d})
    rec = ngpu.CreateRecord("", ["spike_height"] * 6, list(range(det[0], det[0] + 6)), [0] * 6)
    ngpu.Simulate(12.0)
    data = np.array(ngpu.GetRecordData(rec), dtype=np.float32)
    for (k, j, w, d) in conns:
        expect = np.zeros(data.shape[0], np.float32)
        for ts, hg in zip(times[k], heights[k]):
            row = int(round(ts / 0.1)) + 1 + (int(round(d / 0.1)) - 1) + 1 + 1
            expect[row] += np.float32(w) * np.float32(hg)
        np.testing.assert_array_equal(data[:, 1 + 3 * k + j], expect)


def _oracle_network(net, model, sim_ms, prec=0, poiss_w=1.0, status=None):
    from oracle.oracle import Oracle
    o = Oracle(prec=prec)
    n0 = o.create(model, net["n"], 2)
    for k, v in (status or {}).items():
        if isinstance(v, list):
            for p, vv in enumerate(v):
                o.set(n0, net["n"], k, [vv] * net["n"], port=p)
        else:
            o.set(n0, net["n"], k, [v] * net["n"])
    pg = o.create("poisson_generator", 1, 0)
    o.set(pg, 1, "rate", [net["rate"]])
    o.connect(np.full(net["n"], pg, np.int32), np.arange(net["n"], dtype=np.int32) + n0, 0,
              poiss_w, 0.1)
    se, te, we, de = net["exc"]
    si, ti, wi, di = net["inh"]
    o.connect(se + n0, te + n0, 0, we, de)
    o.connect(si + n0, ti + n0, 1, wi, di)
    o.simulate(sim_ms)
    return o.events()


def _gpu_network(ngpu, net, model, sim_ms, poiss_w=1.0, status=None):
    nrn = ngpu.Create(model, net["n"], 2)
    if status:
        ngpu.SetStatus(nrn, status)
    pg = ngpu.Create("poisson_generator")
    ngpu.SetStatus(pg, "rate", net["rate"])
    # oracle and GPU must give each connection the same stream id: same
    # connection order per source (poisson first, one source)
    ngpu.Connect(pg, nrn, {"rule": "all_to_all"}, {"weight": poiss_w, "delay": 0.1, "receptor": 0})
    se, te, we, de = net["exc"]
    si, ti, wi, di = net["inh"]
    netbuild.connect_edges_fast(ngpu, se + nrn[0], te + nrn[0], 0, we, de)
    netbuild.connect_edges_fast(ngpu, si + nrn[0], ti + nrn[0], 1, wi, di)
    ngpu.ActivateRecSpikeTimes(nrn, 5000)
    ngpu.Simulate(sim_ms)
    st = ngpu.GetRecSpikeTimes(nrn)
    t_idx, node = [], []
    for i, ts in enumerate(st):
        for t in ts:
            t_idx.append(int(round(t / 0.1)))
            node.append(i + nrn[0])
    t_idx = np.array(t_idx, np.int64)
    node = np.array(node, np.int32)
    o = np.lexsort((node, t_idx))
    return t_idx[o], node[o]


@pytest.mark.parametrize("runs", ["1", "0"])
def test_iaf_psc_exp_network_bitexact(ngpu, runs, monkeypatch):
    # iaf_psc_exp + Poisson + delivery: the same float operations in the same
    # order on both sides, so spike rasters must be identical; the shared
    # generator's edges as one uniform run (runs=1) and as per-channel CSR
    monkeypatch.setenv("NGPU_POISSON_RUNS", runs)
    net = netbuild.c1_network(seed=3, n=600, k=60, t_rate=8000.0)
    net["exc"] = (net["exc"][0], net["exc"][1], net["exc"][2] * 60.0, net["exc"][3])
    net["inh"] = (net["inh"][0], net["inh"][1], net["inh"][2] * -60.0, net["inh"][3])
    tg, ng = _gpu_network(ngpu, net, "iaf_psc_exp", 300.0, poiss_w=30.0)
    to, no = _oracle_network(net, "iaf_psc_exp", 300.0, poiss_w=30.0)
    print("iaf spikes gpu", len(tg), "oracle", len(to))
    assert len(to) > 500
    np.testing.assert_array_equal(tg, to)
    np.testing.assert_array_equal(ng, no)


@pytest.mark.parametrize("rows", [None, "2", "5"])
def test_iaf_psc_exp_network_tiled_bitexact(ngpu, monkeypatch, rows):
    # the same network with the dense tiled delivery (LDS tile sums added to
    # the ring once per step) instead of global atomics: identical raster;
    # rows = the delay rows kept in LDS (default: all but 1% of the synapses;
    # 2 and 5 send most records to the ring by global atomics)
    monkeypatch.setenv("NGPU_TILED", "1")
    if rows:
        monkeypatch.setenv("NGPU_TILE_ROWS", rows)
    net = netbuild.c1_network(seed=3, n=600, k=60, t_rate=8000.0)
    net["exc"] = (net["exc"][0], net["exc"][1], net["exc"][2] * 60.0, net["exc"][3])
    net["inh"] = (net["inh"][0], net["inh"][1], net["inh"][2] * -60.0, net["inh"][3])
    tg, ng = _gpu_network(ngpu, net, "iaf_psc_exp", 300.0, poiss_w=30.0)
    cfg = ngpu.GetBatchConfig()
    to, no = _oracle_network(net, "iaf_psc_exp", 300.0, poiss_w=30.0)
    print("iaf tiled spikes gpu", len(tg), "oracle", len(to), cfg)
    assert cfg["tiled"] == 1
    if rows:
        assert cfg["tile_rows"] == int(rows)
    assert len(to) > 500
    np.testing.assert_array_equal(tg, to)
    np.testing.assert_array_equal(ng, no)


@pytest.mark.parametrize("rows", [None, "2"])
def test_tiled_delivery_rk5_split_steps(ngpu, rows):
    # one RK5 population with split steps (lagged spikes, their d_int = 0
    # synapses one slot earlier): the tiled delivery launch after every step
    # against fused delivery inside the next launch: identical rasters, state
    # and attempt counts, over two Simulate calls (rows 2: only the slots
    # t and t + 1 in LDS, the rest by global atomics)
    net = netbuild.c1_network(seed=7)
    res = []
    for tiled in ("1", "0"):
        os.environ["NGPU_TILED"] = tiled
        os.environ["NGPU_RK5_DEFER"] = "3"
        if rows:
            os.environ["NGPU_TILE_ROWS"] = rows
        try:
            ngpu.ResetKernel()
            ngpu.SetVerbosityLevel(0)
            t, n = _gpu_network(ngpu, net, "aeif_cond_beta_multisynapse", 60.0,
                                poiss_w=netbuild.C1_POISSON_W, status=netbuild.C1_RECEPTORS)
            cfg = ngpu.GetBatchConfig()
            nodes = ngpu.NodeSeq(0, net["n"])
            ngpu.Simulate(20.0)
            v = np.array(ngpu.GetStatus(nodes, "V_m"))
            st = ngpu.GetRecSpikeTimes(nodes)
            res.append((t, n, v, [list(x) for x in st], ngpu.GetStepStats(), cfg))
        finally:
            os.environ.pop("NGPU_TILED", None)
            os.environ.pop("NGPU_RK5_DEFER", None)
            os.environ.pop("NGPU_TILE_ROWS", None)
    (t1, n1, v1, st1, s1, c1), (t0, n0, v0, st0, s0, c0) = res
    print("tiled", c1, "split steps", s1["split_steps"], "spikes", len(t1))
    assert c1["tiled"] == 1 and c0["tiled"] == 0
    assert s1["split_steps"] > 300 and sum(len(x) for x in st1) > 50
    np.testing.assert_array_equal(t1, t0)
    np.testing.assert_array_equal(n1, n0)
    np.testing.assert_array_equal(v1, v0)
    assert st1 == st0
    assert s1["attempts"] == s0["attempts"] and s1["syn_events"] == s0["syn_events"]


def _iaf_one_to_one_poisson(ngpu, net, rates):
    nrn = ngpu.Create("iaf_psc_exp", net["n"], 2)
    pg = ngpu.Create("poisson_generator", net["n"])
    ngpu.SetStatus(pg, {"rate": {"array": [float(r) for r in rates]}})
    ngpu.Connect(pg, nrn, {"rule": "one_to_one"}, {"weight": 30.0, "delay": 0.1, "receptor": 0})
    se, te, we, de = net["exc"]
    si, ti, wi, di = net["inh"]
    netbuild.connect_edges_fast(ngpu, se + nrn[0], te + nrn[0], 0, we * 60.0, de)
    netbuild.connect_edges_fast(ngpu, si + nrn[0], ti + nrn[0], 1, wi * -60.0, di)
    ngpu.ActivateRecSpikeTimes(nrn, 5000)
    ngpu.Simulate(200.0)
    return [list(ts) for ts in ngpu.GetRecSpikeTimes(nrn)]


def test_poisson_one_to_one_run_matches_csr(ngpu, monkeypatch):
    # one generator per neuron (generator index stepping with the channel):
    # the uniform-run fold and the per-channel CSR fold must give the same
    # raster bit for bit (the CSR path is pinned to the oracle above)
    net = netbuild.c1_network(seed=5, n=400, k=40, t_rate=8000.0)
    rates = np.random.default_rng(9).uniform(4000.0, 12000.0, net["n"])
    out = {}
    for runs in ("1", "0"):
        monkeypatch.setenv("NGPU_POISSON_RUNS", runs)
        ngpu.ResetKernel()
        ngpu.SetVerbosityLevel(0)
        out[runs] = _iaf_one_to_one_poisson(ngpu, net, rates)
    n_spk = sum(len(t) for t in out["1"])
    print("one-to-one poisson spikes", n_spk)
    assert n_spk > 100
    assert out["1"] == out["0"]


def test_aeif_network_config1(ngpu):
    # config-1 network (1000 AdEx, 2 ports, K=100, 20 kHz Poisson), 100 ms:
    # spike times must match the oracle over the first 100 ms
    net = netbuild.c1_network(seed=1)
    status = netbuild.C1_RECEPTORS
    tg, ng = _gpu_network(ngpu, net, "aeif_cond_beta_multisynapse", 100.0, poiss_w=netbuild.C1_POISSON_W,
                          status=status)
    to, no = _oracle_network(net, "aeif_cond_beta_multisynapse", 100.0, poiss_w=netbuild.C1_POISSON_W,
                             status=status)
    print("aeif spikes gpu", len(tg), "oracle", len(to))
    assert len(to) > 100
    a = set(zip(tg.tolist(), ng.tolist()))
    b = set(zip(to.tolist(), no.tolist()))
    print("aeif mismatched spikes", len(a ^ b))
    assert a == b


def test_spike_count_and_times_consistency(ngpu):
    # python/test/test_spike_times.py:8-124: spike_count, detector rows and
    # recorded spike times agree, across two Simulate calls with buffering
    n_neurons = 30
    pg = ngpu.Create("poisson_generator")
    ngpu.SetStatus(pg, "rate", 500.0)
    neuron = ngpu.Create("aeif_cond_beta_multisynapse", n_neurons)
    ngpu.ActivateSpikeCount(neuron)
    ngpu.ActivateRecSpikeTimes(neuron, 500)
    ngpu.SetRecSpikeTimesStep(neuron, 5)
    sd = ngpu.Create("spike_detector", n_neurons)
    ngpu.Connect(pg, neuron, {"rule": "all_to_all"}, {"weight": 4.0, "delay": 0.2})
    ngpu.Connect(neuron, sd, {"rule": "one_to_one"}, {"weight": 1.0, "delay": 0.1})
    record = ngpu.CreateRecord("", ["spike_height"] * n_neurons, sd.ToList(), [0] * n_neurons)
    ngpu.Simulate(490)
    ngpu.SetStatus(pg, "rate", 0.0)
    ngpu.Simulate(10)
    data = ngpu.GetRecordData(record)
    spike = np.sum(np.array(data)[:, 1:], axis=0)
    spike_times = [[] for _ in range(n_neurons)]
    for row in data:
        for i in range(1, len(row)):
            if row[i] > 0.5:
                spike_times[i - 1].append(round(row[0] - 0.2, 4))
    spike_count = ngpu.GetStatus(neuron, "spike_count")
    assert sum(c[0] for c in spike_count) > 0
    for i in range(n_neurons):
        assert abs(spike[i] - spike_count[i][0]) < 1e-6
    rec_times = ngpu.GetRecSpikeTimes(neuron)
    for j in range(n_neurons):
        assert len(rec_times[j]) == spike_count[j][0]
        for i, t in enumerate(rec_times[j]):
            assert abs(round(t, 4) - spike_times[j][i]) < 1e-6


def test_two_simulate_calls_continue(ngpu):
    sc = SCENARIOS["aeif_cond_beta_multisynapse"]
    from oracle.oracle import Oracle
    n = ngpu.Create("aeif_cond_beta_multisynapse", 1, 1)
    ngpu.SetStatus(n, {"I_e": 600.0})
    rec = ngpu.CreateRecord("", ["V_m"], [n[0]], [0])
    ngpu.Simulate(30.0)
    ngpu.Simulate(45.0)
    v = np.array([r[1] for r in ngpu.GetRecordData(rec)])
    o = Oracle(prec=0)
    m = o.create("aeif_cond_beta_multisynapse", 1, 1)
    o.set(m, 1, "I_e", [600.0])
    r = o.record(m, "V_m")
    o.simulate(30.0)
    o.simulate(45.0)
    vo = o.record_data(r)[1]
    assert len(v) == len(vo) == 751
    assert np.max(np.abs(v - vo) / np.abs(vo)) < 1e-5


def test_persistent_rk5_matches_lane_per_neuron_kernel(ngpu):
    # the persistent (lane-refill) RK5 kernel and the one-lane-per-neuron
    # kernel run the same per-neuron arithmetic: identical rasters and state
    net = netbuild.c1_network(seed=7)
    res = []
    for mode in ("1", "0"):
        os.environ["NGPU_RK5_PERSISTENT"] = mode
        try:
            ngpu.ResetKernel()
            ngpu.SetVerbosityLevel(0)
            t, n = _gpu_network(ngpu, net, "aeif_cond_beta_multisynapse", 60.0,
                                poiss_w=netbuild.C1_POISSON_W, status=netbuild.C1_RECEPTORS)
            v = np.array(ngpu.GetStatus(ngpu.NodeSeq(0, net["n"]), "V_m"))
            res.append((t, n, v, ngpu.GetStepStats()))
        finally:
            os.environ.pop("NGPU_RK5_PERSISTENT", None)
    (t1, n1, v1, s1), (t0, n0, v0, s0) = res
    assert len(t1) > 50
    np.testing.assert_array_equal(t1, t0)
    np.testing.assert_array_equal(n1, n0)
    np.testing.assert_array_equal(v1, v0)
    assert s1["attempts"] == s0["attempts"] and s1["rejected"] == s0["rejected"]


def _mixed_gpu(ngpu, net, sim_ms):
    pops = []
    for p in net["pops"]:
        nrn = ngpu.Create(p["model"], p["n"], p["n_port"])
        assert nrn[0] == p["i0"]
        if p["status"]:
            ngpu.SetStatus(nrn, p["status"])
        pops.append(nrn)
    for p, nrn in zip(net["pops"], pops):
        pg = ngpu.Create("poisson_generator")
        ngpu.SetStatus(pg, "rate", p["rate"])
        ngpu.Connect(pg, nrn, {"rule": "all_to_all"},
                     {"weight": p["pw"], "delay": 0.1, "receptor": 0})
    for (s, t, port, w, d) in net["edges"]:
        for q in (0, 1):
            m = port == q
            netbuild.connect_edges_fast(ngpu, s[m], t[m], q, w[m], d[m])
    for nrn in pops:
        ngpu.ActivateRecSpikeTimes(nrn, 5000)
    ngpu.Simulate(sim_ms)
    t_idx, node = [], []
    for nrn in pops:
        for i, ts in enumerate(ngpu.GetRecSpikeTimes(nrn)):
            for tt in ts:
                t_idx.append(int(round(tt / 0.1)))
                node.append(i + nrn[0])
    t_idx = np.array(t_idx, np.int64)
    node = np.array(node, np.int32)
    o = np.lexsort((node, t_idx))
    return t_idx[o], node[o]


def _mixed_oracle(net, sim_ms):
    from oracle.oracle import Oracle
    o = Oracle(prec=0)
    for p in net["pops"]:
        n0 = o.create(p["model"], p["n"], p["n_port"])
        assert n0 == p["i0"]
        for k, v in p["status"].items():
            for q, vv in enumerate(v):
                o.set(n0, p["n"], k, [vv] * p["n"], port=q)
    for p in net["pops"]:
        pg = o.create("poisson_generator", 1, 0)
        o.set(pg, 1, "rate", [p["rate"]])
        o.connect(np.full(p["n"], pg, np.int32), np.arange(p["n"], dtype=np.int32) + p["i0"], 0,
                  p["pw"], 0.1)
    for (s, t, port, w, d) in net["edges"]:
        for q in (0, 1):
            m = port == q
            o.connect(s[m], t[m], q, w[m], d[m])
    o.simulate(sim_ms)
    return o.events()


@pytest.mark.parametrize("split", ["default", "2"])
def test_mixed_model_network(ngpu, split):
    # six populations of five models (izhikevich_cond_beta x2, aeif_psc_exp,
    # aeif_cond_alpha_multisynapse with 4 ports, aeif_cond_beta_multisynapse,
    # iaf_psc_exp), fixed indegree 50 across them, delays 0.1-2 ms: the
    # raster over 100 ms equals the oracle's spike for spike.  "2" splits
    # every neuron-step longer than 2 attempts over two launches.
    net = netbuild.mixed_network()
    if split != "default":
        os.environ["NGPU_RK5_DEFER"] = split
    try:
        ngpu.ResetKernel()
        ngpu.SetVerbosityLevel(0)
        tg, ng = _mixed_gpu(ngpu, net, 100.0)
        st = ngpu.GetStepStats()
    finally:
        os.environ.pop("NGPU_RK5_DEFER", None)
    to, no = _mixed_oracle(net, 100.0)
    print("mixed spikes gpu", len(tg), "oracle", len(to), "split steps", st["split_steps"])
    for p in net["pops"]:
        m = (no >= p["i0"]) & (no < p["i0"] + p["n"])
        assert m.sum() > 20, p["model"]
    if split != "default":
        assert st["split_steps"] > 1000
    np.testing.assert_array_equal(tg, to)
    np.testing.assert_array_equal(ng, no)


def _mixed_run_calls(ngpu, net, calls, batch_cfg=None, autotune_after=None):
    """The mixed network over consecutive Simulate calls of the given lengths
    (ms); optional batch configuration, and an autotuner run (steps per
    candidate) after the first call.  Returns raster, final V of every
    population and the step counters."""
    pops = []
    for p in net["pops"]:
        nrn = ngpu.Create(p["model"], p["n"], p["n_port"])
        if p["status"]:
            ngpu.SetStatus(nrn, p["status"])
        pops.append(nrn)
    for p, nrn in zip(net["pops"], pops):
        pg = ngpu.Create("poisson_generator")
        ngpu.SetStatus(pg, "rate", p["rate"])
        ngpu.Connect(pg, nrn, {"rule": "all_to_all"},
                     {"weight": p["pw"], "delay": 0.1, "receptor": 0})
    for (s, t, port, w, d) in net["edges"]:
        for q in (0, 1):
            m = port == q
            netbuild.connect_edges_fast(ngpu, s[m], t[m], q, w[m], d[m])
    for nrn in pops:
        ngpu.ActivateRecSpikeTimes(nrn, 5000)
    ngpu.Calibrate()
    if batch_cfg is not None:
        ngpu.SetBatchConfig(**batch_cfg)
    cfg = ngpu.GetBatchConfig()
    tuned = None
    for k, ms in enumerate(calls):
        ngpu.Simulate(ms)
        if k == 0 and autotune_after:
            tuned = ngpu.AutotuneBatch(autotune_after)
    t_idx, node = [], []
    for nrn in pops:
        for i, ts in enumerate(ngpu.GetRecSpikeTimes(nrn)):
            for tt in ts:
                t_idx.append(int(round(tt / 0.1)))
                node.append(i + nrn[0])
    t_idx = np.array(t_idx, np.int64)
    node = np.array(node, np.int32)
    o = np.lexsort((node, t_idx))
    v = [np.array(ngpu.GetStatus(nrn, "V_m_rel" if p["model"] == "iaf_psc_exp" else "V_m"),
                  np.float32).reshape(-1) for p, nrn in zip(net["pops"], pops)]
    return t_idx[o], node[o], np.concatenate(v), ngpu.GetStepStats(), cfg, tuned


def test_batch_kernel_matches_population_launches(ngpu):
    # the six-population mixed network as one batch launch per step (graph
    # replay, default configuration), as direct batch launches with other
    # register / unroll / block-weight settings, and after the autotuner has
    # run on live steps, against one launch per population: rasters, final
    # membrane potentials and RK5 attempt counts are identical (scheduling
    # never changes arithmetic); two or three Simulate calls each
    net = netbuild.mixed_network(seed=11)
    runs = {}
    cases = [("populations", "0", "0", None, None, (50.0, 50.0)),
             ("batch", "1", "0", None, None, (50.0, 50.0)),
             ("batch_direct", "1", "0",
              dict(variant=1, iaf_unroll=4, rk5_weight=2.0, graph=False), None, (50.0, 50.0)),
             ("batch_tiled", "1", "1", None, None, (50.0, 50.0)),
             ("autotuned", "1", "1", None, 4, (30.0, 70.0 - 36 * 4 * 0.1))]
    for name, env, tiled, cfg, tune, calls in cases:
        print("batch case", name, flush=True)
        os.environ["NGPU_BATCH"] = env
        os.environ["NGPU_TILED"] = tiled
        try:
            ngpu.ResetKernel()
            ngpu.SetVerbosityLevel(0)
            runs[name] = _mixed_run_calls(ngpu, net, calls, cfg, tune)
        finally:
            os.environ.pop("NGPU_BATCH", None)
            os.environ.pop("NGPU_TILED", None)
    t0, n0, v0, s0, c0, _ = runs["populations"]
    assert c0["on"] == 0
    assert len(t0) > 500
    for name in ("batch", "batch_direct", "batch_tiled", "autotuned"):
        t, n, v, st, cfg, tuned = runs[name]
        print(name, cfg, "tuned", tuned, "spikes", len(t), "split", st["split_steps"])
        assert cfg["on"] == 1 and cfg["populations"] == 6
        if name == "batch_tiled":
            assert cfg["tiled"] == 1
        np.testing.assert_array_equal(t, t0)
        np.testing.assert_array_equal(n, n0)
        np.testing.assert_array_equal(v, v0)
        assert st["attempts"] == s0["attempts"] and st["rejected"] == s0["rejected"]
        assert st["syn_events"] == s0["syn_events"] and st["spikes"] == s0["spikes"]
    assert runs["autotuned"][5][1] == 36     # with the tiled delivery as a knob


def test_split_steps_match_unsplit(ngpu):
    # a neuron-step needing more attempts than the cap is stored mid-step and
    # finished first in the next launch (its fold input applied then, its
    # spike delivered with its own emission time): rasters, final state and
    # attempt counts equal those of the unsplit run
    net = netbuild.c1_network(seed=7)
    res = []
    for cap in ("3", "0"):
        os.environ["NGPU_RK5_DEFER"] = cap
        try:
            ngpu.ResetKernel()
            ngpu.SetVerbosityLevel(0)
            t, n = _gpu_network(ngpu, net, "aeif_cond_beta_multisynapse", 60.0,
                                poiss_w=netbuild.C1_POISSON_W, status=netbuild.C1_RECEPTORS)
            nodes = ngpu.NodeSeq(0, net["n"])
            v = np.array(ngpu.GetStatus(nodes, "V_m"))
            g1 = np.array(ngpu.GetStatus(nodes, "g1"))
            res.append((t, n, v, g1, ngpu.GetStepStats()))
        finally:
            os.environ.pop("NGPU_RK5_DEFER", None)
    (t1, n1, v1, g11, s1), (t0, n0, v0, g10, s0) = res
    print("split steps", s1["split_steps"], "spikes", len(t1))
    assert s1["split_steps"] > 1000 and s0["split_steps"] == 0
    assert len(t1) > 50
    np.testing.assert_array_equal(t1, t0)
    np.testing.assert_array_equal(n1, n0)
    np.testing.assert_array_equal(v1, v0)
    np.testing.assert_array_equal(g11, g10)
    assert s1["attempts"] == s0["attempts"] and s1["rejected"] == s0["rejected"]


def test_fold_in_rk5_matches_fold_launch(ngpu):
    # the fold of a step done by the next RK5 launch as each lane takes its
    # neuron (default) against a fold launch after every step: with many
    # split steps and over two Simulate() calls (the last step of a call
    # folds by launch, the next call's first launch folds nothing), rasters,
    # state and attempt counts are equal
    net = netbuild.c1_network(seed=9)
    res = []
    for fold_in in ("1", "0"):
        os.environ["NGPU_RK5_DEFER"] = "3"
        os.environ["NGPU_FOLD_IN_RK5"] = fold_in
        try:
            ngpu.ResetKernel()
            ngpu.SetVerbosityLevel(0)
            nrn = ngpu.Create("aeif_cond_beta_multisynapse", net["n"], 2)
            ngpu.SetStatus(nrn, netbuild.C1_RECEPTORS)
            pg = ngpu.Create("poisson_generator")
            ngpu.SetStatus(pg, "rate", net["rate"])
            ngpu.Connect(pg, nrn, {"rule": "all_to_all"},
                         {"weight": netbuild.C1_POISSON_W, "delay": 0.1, "receptor": 0})
            se, te, we, de = net["exc"]
            si, ti, wi, di = net["inh"]
            netbuild.connect_edges_fast(ngpu, se + nrn[0], te + nrn[0], 0, we, de)
            netbuild.connect_edges_fast(ngpu, si + nrn[0], ti + nrn[0], 1, wi, di)
            ngpu.ActivateRecSpikeTimes(nrn, 5000)
            ngpu.Simulate(30.0)
            ngpu.Simulate(30.0)
            st = [list(x) for x in ngpu.GetRecSpikeTimes(nrn)]
            v = np.array(ngpu.GetStatus(nrn, "V_m"))
            g1 = np.array(ngpu.GetStatus(nrn, "g1"))
            res.append((st, v, g1, ngpu.GetStepStats()))
        finally:
            os.environ.pop("NGPU_RK5_DEFER", None)
            os.environ.pop("NGPU_FOLD_IN_RK5", None)
    (st1, v1, g11, s1), (st0, v0, g10, s0) = res
    n_spk = sum(len(x) for x in st1)
    print("split steps", s1["split_steps"], "spikes", n_spk)
    assert s1["split_steps"] > 500 and n_spk > 50
    assert st1 == st0
    np.testing.assert_array_equal(v1, v0)
    np.testing.assert_array_equal(g11, g10)
    assert s1["attempts"] == s0["attempts"] and s1["rejected"] == s0["rejected"]


@pytest.mark.parametrize("cap,d0", [("3", "0"), ("24", "0"), ("3", "1")])
def test_fused_delivery_matches_delivery_launch(ngpu, cap, d0):
    # delivery of a step's spikes by the next RK5 launch's idle waves against
    # a delivery launch after every step, with the d_int = 0 synapses of
    # lagged spikes by a launch after the update launch (d0 = 0, the
    # default) or at the emitting wave's exit (1): with
    # split steps, delays from 0.1 ms and three Simulate() calls (each call's
    # last step delivers by launch, the next call's first launch delivers
    # nothing), rasters, state, attempts and delivered event counts are equal
    net = netbuild.c1_network(seed=13)
    net["exc"] = (net["exc"][0], net["exc"][1], net["exc"][2], np.full_like(net["exc"][3], 0.1))
    res = []
    for fused in ("1", "0"):
        os.environ["NGPU_RK5_DEFER"] = cap
        os.environ["NGPU_FUSED_DELIVER"] = fused
        os.environ["NGPU_LAG_D0"] = d0
        try:
            ngpu.ResetKernel()
            ngpu.SetVerbosityLevel(0)
            nrn = ngpu.Create("aeif_cond_beta_multisynapse", net["n"], 2)
            ngpu.SetStatus(nrn, netbuild.C1_RECEPTORS)
            pg = ngpu.Create("poisson_generator")
            ngpu.SetStatus(pg, "rate", net["rate"])
            ngpu.Connect(pg, nrn, {"rule": "all_to_all"},
                         {"weight": netbuild.C1_POISSON_W, "delay": 0.1, "receptor": 0})
            se, te, we, de = net["exc"]
            si, ti, wi, di = net["inh"]
            netbuild.connect_edges_fast(ngpu, se + nrn[0], te + nrn[0], 0, we, de)
            netbuild.connect_edges_fast(ngpu, si + nrn[0], ti + nrn[0], 1, wi, di)
            ngpu.ActivateRecSpikeTimes(nrn, 5000)
            for ms in (20.0, 25.0, 15.0):
                ngpu.Simulate(ms)
            st = [list(x) for x in ngpu.GetRecSpikeTimes(nrn)]
            v = np.array(ngpu.GetStatus(nrn, "V_m"))
            g1 = np.array(ngpu.GetStatus(nrn, "g1"))
            res.append((st, v, g1, ngpu.GetStepStats()))
        finally:
            os.environ.pop("NGPU_RK5_DEFER", None)
            os.environ.pop("NGPU_FUSED_DELIVER", None)
            os.environ.pop("NGPU_LAG_D0", None)
    (st1, v1, g11, s1), (st0, v0, g10, s0) = res
    n_spk = sum(len(x) for x in st1)
    print("cap", cap, "split steps", s1["split_steps"], "spikes", n_spk, "events", s1["syn_events"])
    assert n_spk > 50 and (cap != "3" or s1["split_steps"] > 500)
    assert st1 == st0
    np.testing.assert_array_equal(v1, v0)
    np.testing.assert_array_equal(g11, g10)
    for k in ("attempts", "rejected", "syn_events", "spikes", "dropped_spikes"):
        assert s1[k] == s0[k], k


@pytest.mark.parametrize("hold", [0.0, 1.0])
def test_parrot_neuron_matches_oracle(ngpu, hold):
    # src/parrot_neuron.cu: the chain generator -> parrot -> detector, records
    # identical to the oracle's (whose timeline tests/test_oracle_pins.py
    # pins by hand)
    from test_oracle_pins import _parrot_chain
    vg, hg = _parrot_chain(hold, ngpu)
    vo, ho = _parrot_chain(hold, None)
    np.testing.assert_array_equal(vg, vo)
    np.testing.assert_array_equal(hg, ho)
    assert np.count_nonzero(hg) == 1


def test_sampled_kernel_timing(ngpu):
    # timing modes 3 / 4 bracket the launches of the steps whose index in the
    # Simulate call is 7 mod 16 only (modes 1 / 2: every step); 2 and 4 the
    # RK5 launches alone
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    n = ngpu.Create("aeif_cond_beta_multisynapse", 100, 2)
    ngpu.SetStatus(n, {"I_e": 600.0})
    ngpu.Simulate(1.0)
    steps = 64
    sampled = len([i for i in range(steps) if i % 16 == 7])
    for mode, expect in ((1, steps), (2, steps), (3, sampled), (4, sampled)):
        ngpu.SetKernelTiming(mode)
        ngpu.Simulate(steps * 0.1)
        kt = ngpu.GetKernelTimes()
        ngpu.SetKernelTiming(False)
        assert kt["rk5_update"][1] == expect, (mode, kt)
        assert kt["rk5_update"][0] > 0.0
        if mode in (2, 4):
            assert all(kt[k][1] == 0 for k in kt if k != "rk5_update"), (mode, kt)
