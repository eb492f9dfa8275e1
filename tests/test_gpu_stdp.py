"""Plastic synapses: the reference's STDP known answer
(python/test/test_stdp.py).  Fifty pairs of aeif_cond_beta_multisynapse
neurons fire once each: the pre-synaptic neuron at ~51 ms, the
post-synaptic one at ~451 ms, and the STDP synapse between them has delay
400 ms - dt_i, so the pre-synaptic spike arrives dt_i before the
post-synaptic spike (dt_i = -122.5 ... 122.5 ms in steps of 5 ms).  Each
weight is updated exactly once -- potentiated at the post-synaptic spike
when dt_i > 0 (src/rev_spike.cu:40-56), depressed at the arrival when
dt_i < 0 (src/get_spike.cu:80-95) -- by the rule of src/stdp.cu:36-61,
restated in closed form below as the reference's test states it."""
import math

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ngpu():
    if not gpu_available():
        pytest.skip("no GPU")
    import nestgpu
    return nestgpu


def stdp_closed_form(w, dt, tau_plus, tau_minus, w_plus, alpha, mu_plus, mu_minus, w_max):
    if dt >= 0:
        w1 = w + w_plus * math.exp(-dt / tau_plus) * math.pow(1.0 - w / w_max, mu_plus)
        return min(w1, w_max)
    w1 = w - alpha * w_plus * math.exp(dt / tau_minus) * math.pow(w / w_max, mu_minus)
    return max(w1, 0.0)


def _check_weights(ngpu, w, w0, dts, pre, post, p, time_diff=400.0):
    """Each weight equals the closed form at its pair's actual spike-time
    difference: post-synaptic spike time minus (pre-synaptic spike time +
    delay), both recorded.  That is dt_i itself unless a neuron crossed
    threshold a step or two later (the plastic synapse's own small input
    changes the adaptive step sizes of its target); the reference's 1e-5
    tolerance against dt_i absorbs those shifts too."""
    args = (p["tau_plus"], p["tau_minus"], p["lambda"] * p["Wmax"], p["alpha"], p["mu_plus"],
            p["mu_minus"], p["Wmax"])
    # spike times are read for a whole node group (pre and post together)
    t_all = ngpu.GetRecSpikeTimes(ngpu.NodeSeq(pre.i0, pre.n + post.n))
    t_pre, t_post = t_all[:pre.n], t_all[pre.n:]
    nominal = 0
    for i, dt in enumerate(dts):
        assert len(t_pre[i]) == 1 and len(t_post[i]) == 1, (i, t_pre[i], t_post[i])
        dt_actual = round((t_post[i][0] - (t_pre[i][0] + time_diff - dt)) * 10.0) / 10.0
        nominal += abs(dt_actual - dt) < 1e-6
        expect = stdp_closed_form(w0, dt_actual, *args)
        assert abs(w[i] - expect) <= 2e-6 * expect, (i, dt, dt_actual, w[i], expect)
        assert abs(w[i] - stdp_closed_form(w0, dt, *args)) < 1e-5
    print("pairs at the nominal dt: %d of %d" % (nominal, len(dts)))
    assert nominal >= len(dts) // 2


def _stdp_network(ngpu, n, params, time_diff=400.0, dt_step=5.0):
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    sg_syn = ngpu.CreateSynGroup("stdp", params)
    sg = ngpu.Create("spike_generator", n)
    neuron = ngpu.Create("aeif_cond_beta_multisynapse", 2 * n)
    ngpu.SetStatus(neuron, {"t_ref": 1000.0, "den_delay": 0.0})
    ngpu.ActivateRecSpikeTimes(neuron, 10)
    pre, post = neuron[0:n], neuron[n:2 * n]
    dts = [dt_step * (-0.5 * (n - 1) + i) for i in range(n)]
    ngpu.SetStatus(sg, {"spike_times": [50.0], "spike_heights": [1.0]})
    ngpu.Connect(sg, pre, {"rule": "one_to_one"}, {"weight": 17.9, "delay": 1.0})
    ngpu.Connect(sg, post, {"rule": "one_to_one"}, {"weight": 17.9, "delay": 1.0 + time_diff})
    w0 = params["Wmax"] / 2
    for i in range(n):
        ngpu.Connect([pre[i]], [post[i]], {"rule": "one_to_one"},
                     {"weight": w0, "delay": time_diff - dts[i], "synapse_group": sg_syn})
    return pre, post, dts, w0, sg_syn


def test_stdp_known_answer(ngpu):
    p = {"tau_plus": 20.0, "tau_minus": 20.0, "lambda": 1.0, "alpha": 1.0, "mu_plus": 1.0,
         "mu_minus": 1.0, "Wmax": 0.001}
    n = 50
    pre, post, dts, w0, sg_syn = _stdp_network(ngpu, n, p)
    ngpu.Simulate(1000.0)
    conn = ngpu.GetConnections(pre, post)
    w = np.array(ngpu.GetStatus(conn, "weight"), dtype=np.float64)
    assert len(w) == n
    expect = np.array([stdp_closed_form(w0, dts[i], p["tau_plus"], p["tau_minus"],
                                        p["lambda"] * p["Wmax"], p["alpha"], p["mu_plus"],
                                        p["mu_minus"], p["Wmax"]) for i in range(n)])
    print("weights", w[:3], "...", w[-3:], "max |dw|", np.abs(w - expect).max())
    # the reference's tolerance (1e-5), and far tighter per weight: the
    # update is the same double formula, rounded to float once
    assert np.abs(w - expect).max() < 1e-5
    _check_weights(ngpu, w, w0, dts, pre, post, p)
    # every weight moved: potentiated for dt > 0, depressed for dt < 0
    assert np.all((w > w0) == (np.array(dts) > 0))
    # the plastic synapses are the only ones of their group
    assert len(ngpu.GetConnections(syn_group=sg_syn.i_syn_group)) == n


def test_stdp_asymmetric_parameters(ngpu):
    # weight-dependent exponents and asymmetric windows: the same closed form
    p = {"tau_plus": 15.0, "tau_minus": 30.0, "lambda": 0.5, "alpha": 0.7, "mu_plus": 0.4,
         "mu_minus": 1.6, "Wmax": 0.002}
    n = 30
    pre, post, dts, w0, _ = _stdp_network(ngpu, n, p)
    ngpu.Simulate(1000.0)
    w = np.array(ngpu.GetStatus(ngpu.GetConnections(pre, post), "weight"), dtype=np.float64)
    _check_weights(ngpu, w, w0, dts, pre, post, p)


def test_syn_model_step_order(ngpu):
    # python/test/test_syn_model.py: the linear rule w += offset + fact * Dt
    # (src/test_syn_model.cu:36-41) on zero-weight synapses whose spikes
    # arrive 2, 1 steps before, in the same step as, and 1, 2 steps after
    # the target's spike: depression at arrival only for a target that
    # fired in an earlier step, potentiation at the target's spike also for
    # an arrival in the same step (Dt = 0), so w = offset + fact * dt
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    fact, offset, n, time_diff = 0.2, 0.03, 5, 10.0
    g = ngpu.CreateSynGroup("test_syn_model")
    ngpu.SetSynGroupParam(g, "fact", fact)
    ngpu.SetSynGroupParam(g, "offset", offset)
    sg = ngpu.Create("spike_generator", n)
    neuron = ngpu.Create("aeif_cond_beta_multisynapse", 2 * n)
    ngpu.SetStatus(neuron, {"t_ref": 10.0})
    pre, post = neuron[0:n], neuron[n:2 * n]
    dts = [0.1 * (-0.5 * (n - 1) + i) for i in range(n)]
    ngpu.SetStatus(sg, {"spike_times": [50.0], "spike_heights": [1.0]})
    ngpu.Connect(sg, pre, {"rule": "one_to_one"},
                 {"weight": 17.9, "delay": 1.0, "receptor": 0, "synapse_group": 0})
    ngpu.Connect(sg, post, {"rule": "one_to_one"},
                 {"weight": 17.9, "delay": 1.0 + time_diff, "receptor": 0, "synapse_group": 0})
    for i in range(n):
        ngpu.Connect([pre[i]], [post[i]], {"rule": "one_to_one"},
                     {"weight": 0.0, "delay": time_diff - dts[i], "receptor": 0, "synapse_group": g})
    ngpu.Simulate(200.0)
    w = np.array(ngpu.GetStatus(ngpu.GetConnections(pre, post), "weight"), dtype=np.float64)
    expect = np.array([dt * fact + offset for dt in dts])
    print("test_syn_model weights", w, "expected", expect)
    np.testing.assert_allclose(w, expect, rtol=0, atol=1e-6)
