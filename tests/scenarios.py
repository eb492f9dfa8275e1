"""Backend-neutral restatements of the reference's single-neuron trace tests.

Each scenario follows one of the reference scripts under python/test/
(cited per scenario).  A backend implements create/set/spike_times/connect/
record/simulate/record_data; `OracleBackend` drives the CPU oracle and
`NestGpuBackend` drives this build's `nestgpu` module, so the same scenario
runs on both and the traces are compared.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def nest_trace(name):
    d = np.loadtxt(os.path.join(GOLDEN, name), delimiter="\t")
    return d[:, 0], d[:, 1]


def trace_rmse(v_gpu, v_nest):
    """The reference metric: std(dV)/|mean(V)| with V sampled at row 10*i+20
    (python/test/test_aeif_cond_beta_multisynapse.py:44-47)."""
    v_gpu = np.asarray(v_gpu, dtype=np.float64)
    dv = np.array([v_gpu[i * 10 + 20] - v_nest[i] for i in range(len(v_nest))])
    return np.std(dv) / abs(np.mean(v_gpu))


# name -> (model, n_port, status, receptors [(port, weight, delay)], var,
#          v_offset, golden file, tolerance)
SCENARIOS = {
    # python/test/test_aeif_cond_beta_multisynapse.py:4-26
    "aeif_cond_beta_multisynapse": dict(
        model="aeif_cond_beta_multisynapse", n_port=3,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0},
        port={"E_rev": [20.0, 0.0, -85.0], "tau_decay": [40.0, 20.0, 30.0],
              "tau_rise": [20.0, 10.0, 5.0]},
        syn=[(0, 0.1, 1.0), (1, 0.2, 100.0), (2, 0.5, 130.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_cond_beta_multisynapse_nest.txt", tol=5e-4),
    # python/test/test_aeif_cond_alpha_multisynapse.py:4-24
    "aeif_cond_alpha_multisynapse": dict(
        model="aeif_cond_alpha_multisynapse", n_port=3,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0},
        port={"E_rev": [20.0, 0.0, -85.0], "tau_syn": [40.0, 20.0, 30.0]},
        syn=[(0, 0.1, 1.0), (1, 0.2, 100.0), (2, 0.5, 130.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_cond_alpha_multisynapse_nest.txt", tol=5e-4),
    # python/test/test_aeif_psc_exp.py:4-28
    "aeif_psc_exp": dict(
        model="aeif_psc_exp", n_port=2,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0,
              "tau_syn_ex": 40.0, "tau_syn_in": 20.0},
        port={},
        syn=[(0, 1.0, 1.0), (1, 2.0, 100.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_psc_exp_nest.txt", tol=5e-6),
    # python/test/test_aeif_cond_alpha.py:4-23
    "aeif_cond_alpha": dict(
        model="aeif_cond_alpha", n_port=2,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0,
              "E_rev_ex": 20.0, "E_rev_in": -85.0, "tau_syn_ex": 40.0, "tau_syn_in": 20.0},
        port={},
        syn=[(0, 0.1, 1.0), (1, 0.2, 100.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_cond_alpha_nest.txt", tol=5e-5),
    # python/test/test_aeif_cond_beta.py:4-25
    "aeif_cond_beta": dict(
        model="aeif_cond_beta", n_port=2,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0,
              "E_rev_ex": 20.0, "E_rev_in": -85.0, "tau_decay_ex": 40.0, "tau_decay_in": 20.0,
              "tau_rise_ex": 20.0, "tau_rise_in": 5.0},
        port={},
        syn=[(0, 0.1, 1.0), (1, 0.2, 100.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_cond_beta_nest.txt", tol=5e-5),
    # python/test/test_aeif_psc_alpha.py:4-26
    "aeif_psc_alpha": dict(
        model="aeif_psc_alpha", n_port=2,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0,
              "tau_syn_ex": 40.0, "tau_syn_in": 20.0},
        port={},
        syn=[(0, 1.0, 1.0), (1, 2.0, 100.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_psc_alpha_nest.txt", tol=3e-6),
    # python/test/test_aeif_psc_alpha_multisynapse.py:4-25
    "aeif_psc_alpha_multisynapse": dict(
        model="aeif_psc_alpha_multisynapse", n_port=2,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0},
        port={"tau_syn": [40.0, 20.0]},
        syn=[(0, 1.0, 1.0), (1, -2.0, 100.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_psc_alpha_multisynapse_nest.txt", tol=5e-4),
    # python/test/test_aeif_psc_delta.py:4-24 (one port: both inputs on port 0)
    "aeif_psc_delta": dict(
        model="aeif_psc_delta", n_port=1,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0, "C_m": 20000.0},
        port={},
        syn=[(0, 1.0, 1.0), (0, -2.0, 100.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_psc_delta_nest.txt", tol=5e-4),
    # python/test/test_aeif_psc_exp_multisynapse.py:4-25
    "aeif_psc_exp_multisynapse": dict(
        model="aeif_psc_exp_multisynapse", n_port=2,
        scal={"V_peak": 0.0, "a": 4.0, "b": 80.5, "E_L": -70.6, "g_L": 300.0},
        port={"tau_syn": [40.0, 20.0]},
        syn=[(0, 1.0, 1.0), (1, -2.0, 100.0)],
        var="V_m", v_offset=0.0,
        golden="test_aeif_psc_exp_multisynapse_nest.txt", tol=5e-4),
    # python/test/test_iaf_psc_exp.py:4-30 (records V_m_rel, adds E_L=-70)
    "iaf_psc_exp": dict(
        model="iaf_psc_exp", n_port=2, scal={}, port={},
        syn=[(0, 1.0, 1.0), (1, -2.0, 100.0)],
        var="V_m_rel", v_offset=-70.0,
        golden="test_iaf_psc_exp_nest.txt", tol=5e-5),
}

SPIKE_TIMES = [10.0, 400.0]
SIM_TIME = 800.0


class OracleBackend:
    def __init__(self, prec=0):
        from oracle.oracle import Oracle
        self.o = Oracle(prec=prec)

    def run_single(self, sc):
        o = self.o
        n = o.create(sc["model"], 1, sc["n_port"])
        for k, v in sc["scal"].items():
            o.set(n, 1, k, [v])
        for k, vals in sc["port"].items():
            for p, v in enumerate(vals):
                o.set(n, 1, k, [v], port=p)
        sg = o.create("spike_generator", 1, 0)
        o.set_spike_times(sg, SPIKE_TIMES)
        for (port, w, d) in sc["syn"]:
            o.connect([sg], [n], [port], [w], [d])
        rec = o.record(n, sc["var"], 0)
        o.simulate(SIM_TIME)
        t, v = o.record_data(rec)
        return t, v.astype(np.float64) + sc["v_offset"]


class NestGpuBackend:
    """Drives the product's Python API exactly as the reference scripts do."""

    def __init__(self, ngpu):
        self.ngpu = ngpu

    def run_single(self, sc):
        ngpu = self.ngpu
        neuron = ngpu.Create(sc["model"], 1, sc["n_port"])
        status = dict(sc["scal"])
        status.update(sc["port"])
        ngpu.SetStatus(neuron, status)
        spike = ngpu.Create("spike_generator")
        ngpu.SetStatus(spike, {"spike_times": SPIKE_TIMES})
        for (port, w, d) in sc["syn"]:
            ngpu.Connect(spike, neuron, {"rule": "all_to_all"},
                         {"receptor": port, "weight": w, "delay": d})
        record = ngpu.CreateRecord("", [sc["var"]], [neuron[0]], [0])
        ngpu.Simulate(SIM_TIME)
        data = ngpu.GetRecordData(record)
        t = np.array([r[0] for r in data], np.float32)
        v = np.array([r[1] for r in data], np.float64) + sc["v_offset"]
        return t, v
