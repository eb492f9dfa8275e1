"""Host-side tests of the C ABI and the nestgpu module (no GPU needed).

Everything before calibration (node creation, status dictionaries,
connection store, connection queries, kernel parameters, error convention)
runs on the host staging path, so it is tested here against the reference's
golden stdout python/test/logp3_connect.txt and its documented behaviour.
"""
import ctypes
import io
import os
import re
import subprocess
import sys
from contextlib import redirect_stdout

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nest-gpu_amd")
LIB = os.path.join(PKG, "libnestgpu_amd.so")
HEADER = os.path.join(ROOT, "include", "nestgpu_C.h")


@pytest.fixture(scope="module")
def ngpu():
    import nestgpu
    return nestgpu


@pytest.fixture(autouse=True)
def fresh(ngpu):
    ngpu.ResetKernel()
    yield


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(NESTGPU_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 133
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_reference_python_bound_symbols_are_exported():
    # the 124 functions the reference module binds at import
    # (pythonlib/nestgpu.py) must all resolve
    lib = ctypes.CDLL(LIB)
    names = ["NESTGPU_GetErrorMessage", "NESTGPU_GetErrorCode", "NESTGPU_SetOnException",
             "NESTGPU_SetRandomSeed", "NESTGPU_SetTimeResolution", "NESTGPU_GetTimeResolution",
             "NESTGPU_Create", "NESTGPU_CreateRecord", "NESTGPU_GetRecordData",
             "NESTGPU_ConnectSeqSeq", "NESTGPU_GetSeqSeqConnections", "NESTGPU_GetConnectionStatus",
             "NESTGPU_ConnectMpiInit", "NESTGPU_RemoteConnectSeqSeq", "NESTGPU_CreateSynGroup",
             "NESTGPU_GetRecSpikeTimes", "NESTGPU_SetNeuronGroupParam", "NESTGPU_RemoteCreate",
             "NESTGPU_GetIntParamNames", "NESTGPU_SetBoolParam"]
    for n in names:
        assert hasattr(lib, n), n


def _connect_scenario(ngpu):
    # python/test/test_connect.py:1-120, printing the same records
    N = 5
    neuron = ngpu.Create("aeif_cond_beta_multisynapse", 2 * N)
    neuron_even = [neuron[2 * i] for i in range(N)]
    neuron_odd = [neuron[2 * i + 1] for i in range(N)]
    e2o_d, e2o_w, o2e_d, o2e_w = [], [], [], []
    for itgt in range(N):
        ite, ito = 2 * itgt, 2 * itgt + 1
        for isrc in range(N):
            ise, iso = 2 * isrc, 2 * isrc + 1
            e2o_d.append(2.0 * N * ise + ito)
            e2o_w.append(100.0 * (2.0 * N * ise + ito))
            o2e_d.append(2.0 * N * iso + ite)
            o2e_w.append(100.0 * (2.0 * N * iso + ite))
    conn_dict = {"rule": "all_to_all"}
    ngpu.Connect(neuron_even, neuron_odd, conn_dict, {"weight_array": e2o_w, "delay_array": e2o_d})
    ngpu.Connect(neuron_odd, neuron_even, conn_dict, {"weight_array": o2e_w, "delay_array": o2e_d})

    def section(title, ids, keys=None):
        print("########################################")
        print(title)
        st = ngpu.GetStatus(ids) if keys is None else ngpu.GetStatus(ids, keys)
        for s in st:
            print(s)
        print()
        print()

    conn_id = ngpu.GetConnections(neuron_even, neuron)
    section("Even to all", conn_id)
    section("Even to all weight, delat", conn_id, ["weight", "delay"])
    section("All to odd", ngpu.GetConnections(neuron, neuron_odd))
    neuron_3_6 = neuron[3:7]
    section("Even to 3,4,5,6", ngpu.GetConnections(neuron_even, neuron_3_6))
    section("3,4,5,6 to odd", ngpu.GetConnections(neuron_3_6, neuron_odd))


def test_connect_golden_stdout(ngpu):
    buf = io.StringIO()
    with redirect_stdout(buf):
        _connect_scenario(ngpu)
    got = buf.getvalue().strip().splitlines()
    gold = open(os.path.join(ROOT, "tests", "golden", "logp3_connect.txt")).read().splitlines()
    start = gold.index("########################################")
    gold = [l for l in gold[start:]]
    while gold and gold[-1] == "":
        gold.pop()
    assert got == gold


def test_status_defaults_and_set(ngpu):
    n = ngpu.Create("aeif_cond_beta_multisynapse", 4, 2)
    st = ngpu.GetStatus(n[0])
    assert st["V_th"] == pytest.approx(-50.4, rel=1e-6)
    assert st["V_m"] == pytest.approx(-70.6, rel=1e-6)
    assert st["tau_decay"] == [20.0, 20.0] and st["tau_rise"] == [2.0, 2.0]
    assert st["h_min_rel"] == pytest.approx(1e-3) and st["h0_rel"] == pytest.approx(1e-2)
    ngpu.SetStatus(n, {"E_rev": [0.0, -85.0], "I_e": 10.0})
    ngpu.SetStatus(n[1:2], {"I_e": 20.0})
    assert ngpu.GetStatus(n, "I_e") == [10.0, 20.0, 10.0, 10.0]
    assert ngpu.GetStatus(n, "E_rev")[3] == [0.0, -85.0]
    ngpu.SetStatus([n[0], n[3]], "V_m", -65.0)
    assert ngpu.GetStatus(n, "V_m") == [-65.0, pytest.approx(-70.6, rel=1e-6),
                                        pytest.approx(-70.6, rel=1e-6), -65.0]
    with pytest.raises(ValueError):
        ngpu.SetStatus(n, {"E_rev": [0.0]})            # wrong port count
    with pytest.raises(ValueError):
        ngpu.SetStatus(n, {"no_such_param": 1.0})


def test_models_and_names(ngpu):
    a = ngpu.Create("iaf_psc_exp", 3)
    b = ngpu.Create("aeif_psc_exp", 2)
    c = ngpu.Create("izhikevich_cond_beta", 2, 2)
    d = ngpu.Create("aeif_cond_alpha_multisynapse", 2, 4)
    assert ngpu.GetScalVarNames(a[0]) == ["I_syn_ex", "I_syn_in", "V_m_rel", "refractory_step"]
    assert ngpu.GetStatus(a[0], "Theta_rel") == 15.0
    assert ngpu.GetScalVarNames(b[0]) == ["I_syn_ex", "I_syn_in", "V_m", "w"]
    assert ngpu.GetStatus(c[0], "u") == pytest.approx(-14.0)
    assert ngpu.GetPortParamNames(d[0]) == ["E_rev", "tau_syn", "g0"]
    with pytest.raises(ValueError, match="outside the scope"):
        ngpu.Create("iaf_psc_alpha", 1)
    with pytest.raises(ValueError, match="Unknown neuron model"):
        ngpu.Create("no_such_model", 1)


def test_kernel_status_and_errors(ngpu):
    ks = ngpu.GetKernelStatus()
    assert set(ks) == {"time_resolution", "max_spike_num_fact", "max_spike_per_host_fact",
                       "rnd_seed", "verbosity_level", "max_spike_buffer_size",
                       "remote_spike_height_flag", "print_time"}
    ngpu.SetKernelStatus({"rnd_seed": 1234, "print_time": True})
    assert ngpu.GetKernelStatus("rnd_seed") == 1234
    with pytest.raises(ValueError, match="MPI is not available"):
        ngpu.MpiId()
    with pytest.raises(ValueError, match="Unknown synapse model name"):
        ngpu.CreateSynGroup("no_such_synapse_model")
    n = ngpu.Create("aeif_cond_beta_multisynapse", 2, 1)
    with pytest.raises(ValueError, match="Delay must be >= time resolution"):
        ngpu.Connect(n, n, {"rule": "all_to_all"}, {"weight": 1.0, "delay": 0.05})
    assert ngpu.GetErrorCode() == 2


def test_fixed_indegree_structure(ngpu):
    src = ngpu.Create("aeif_cond_beta_multisynapse", 200, 2)
    tgt = ngpu.Create("aeif_cond_beta_multisynapse", 50, 2)
    ngpu.Connect(src, tgt, {"rule": "fixed_indegree", "indegree": 30},
                 {"weight": 1.0, "delay": 1.0, "receptor": 1})
    ids = ngpu.GetConnections(src, tgt)
    assert len(ids) == 50 * 30
    per_tgt = {}
    for c in ids:
        s = ngpu.GetStatus(c)
        per_tgt.setdefault(s["target"], []).append(s["source"])
        assert s["port"] == 1
    assert all(len(v) == 30 and len(set(v)) == 30 for v in per_tgt.values())
    # sources differ between targets (independent draws per target)
    assert len({tuple(sorted(v)) for v in per_tgt.values()}) > 40


def test_fixed_outdegree_and_total_number(ngpu):
    a = ngpu.Create("aeif_cond_beta_multisynapse", 20, 1)
    ngpu.Connect(a, a, {"rule": "fixed_outdegree", "outdegree": 7}, {"weight": 1.0, "delay": 0.5})
    for i in range(20):
        t = [ngpu.GetStatus(c, "target") for c in ngpu.GetConnections([a[i]], a)]
        assert len(t) == 7 and len(set(t)) == 7
    ngpu.Connect(a, a, {"rule": "fixed_total_number", "total_num": 33}, {"weight": 2.0, "delay": 0.5})
    w = [ngpu.GetStatus(c, "weight") for c in ngpu.GetConnections(a, a)]
    assert w.count(2.0) == 33


def test_one_to_one_and_delay_groups(ngpu):
    a = ngpu.Create("aeif_cond_beta_multisynapse", 4, 1)
    ngpu.Connect(a, a, {"rule": "one_to_one"}, {"weight": 1.0, "delay": 3.0})
    ngpu.Connect([a[0]], [a[1]], {"rule": "one_to_one"}, {"weight": 2.0, "delay": 1.0})
    ids = ngpu.GetConnections([a[0]])
    st = [ngpu.GetStatus(c) for c in ids]
    # delay groups of a source are delay-ascending (src/connect.cu:56-74)
    assert [s["delay"] for s in st] == [1.0, 3.0]
    assert [(c.i_group, c.i_conn) for c in ids] == [(0, 0), (1, 0)]


def test_normal_clipped_distribution(ngpu):
    import numpy as np
    arr = ngpu.RandomNormalClipped(20000, 1.5, 0.75, 0.1, 3.0, 0.1)
    v = np.array([arr[i] for i in range(20000)])
    assert v.min() >= 0.1 - 1e-6 and v.max() <= 3.0 + 1e-6
    assert np.allclose(np.round((v - 0.1) / 0.1), (v - 0.1) / 0.1, atol=1e-3)


def test_exit_on_error_by_default():
    # with SetOnException(0) a failing call exits with the error code
    code = ("import ctypes; L=ctypes.CDLL(%r); L.NESTGPU_SetOnException(0); "
            "L.NESTGPU_Create(b'no_such_model', 1, 1); print('not reached')" % LIB)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "not reached" not in r.stdout


def test_nodeseq_semantics(ngpu):
    s = ngpu.NodeSeq(10, 5)
    assert len(s) == 5 and s[0] == 10 and s[4] == 14 and s[-1] == 14 and s[-5] == 10
    assert s.ToList() == [10, 11, 12, 13, 14] and list(s) == s.ToList()
    sub = s[1:4]
    assert (sub.i0, sub.n) == (11, 3)
    assert (s.Subseq(0, -1).i0, s.Subseq(0, -1).n) == (10, 5)
    for bad in (lambda: s[5], lambda: s[-6], lambda: s[0:5:2], lambda: s.Subseq(3, 1),
                lambda: s.Subseq(0, 5)):
        with pytest.raises(ValueError):
            bad()
    allseq = ngpu.NodeSeq(None)
    assert (allseq.i0, allseq.n) == (0, -1)


def test_status_shapes_and_dict_values(ngpu):
    n = ngpu.Create("aeif_cond_beta_multisynapse", 3, 2, {"I_e": 5.0})
    assert ngpu.GetStatus(n, "I_e") == [5.0, 5.0, 5.0]
    # a port parameter: one list per node; a group parameter: one value per node
    assert ngpu.GetStatus(n, "E_rev") == [[0.0, 0.0]] * 3
    assert ngpu.GetStatus(n, "h_min_rel") == [pytest.approx(1e-3)] * 3
    assert ngpu.GetStatus([n[0], n[2]], ["I_e", "V_th"]) == [[5.0, pytest.approx(-50.4)]] * 2
    # integer variables read as one-element lists
    ngpu.ActivateSpikeCount(n)
    assert ngpu.GetStatus(n, "spike_count") == [[0], [0], [0]]
    # per-node values from an array dictionary
    ngpu.SetStatus(n, "V_m", {"array": [-60.0, -61.0, -62.0]})
    assert ngpu.GetStatus(n, "V_m") == [-60.0, -61.0, -62.0]
    with pytest.raises(ValueError):
        ngpu.SetStatus(n, "V_m", {"array": [-60.0]})
    with pytest.raises(ValueError):
        ngpu.SetStatus(n, "V_m", {"distribution": "cauchy"})
    # a distribution drawn by the library, clipped to its bounds
    ngpu.SetStatus(n, "I_e", {"distribution": "normal_clipped", "mu": 0.0, "sigma": 5.0,
                              "low": -1.0, "high": 1.0})
    assert all(-1.0 <= v <= 1.0 for v in ngpu.GetStatus(n, "I_e"))
    assert ngpu.GetKernelStatus("print_time") is False
    ngpu.SetKernelStatus("print_time", True)
    assert ngpu.GetKernelStatus("print_time") is True


def test_connect_argument_errors(ngpu):
    a = ngpu.Create("iaf_psc_exp", 4)
    with pytest.raises(ValueError):
        ngpu.Connect(a, a, {"rule": "no_such_rule"}, {"weight": 1.0, "delay": 1.0})
    with pytest.raises(ValueError):
        ngpu.Connect(a, a, {"rule": "all_to_all", "bogus": 1}, {"weight": 1.0, "delay": 1.0})
    with pytest.raises(ValueError):
        ngpu.Connect(a, a, {"rule": "all_to_all"}, {"bogus": 1.0})
    with pytest.raises(ValueError):
        ngpu.Connect(5, a, {"rule": "all_to_all"}, {"weight": 1.0, "delay": 1.0})
    # weights from a distribution dictionary, one per connection
    ngpu.Connect(a, a, {"rule": "all_to_all"},
                 {"weight": {"distribution": "normal_clipped", "mu": 1.0, "sigma": 1.0,
                             "low": 0.5, "high": 1.5}, "delay": 1.0})
    w = [ngpu.GetStatus(c, "weight") for c in ngpu.GetConnections(a, a)]
    assert len(w) == 16 and all(0.5 <= x <= 1.5 for x in w)
    assert ngpu.RuleArraySize({"rule": "fixed_indegree", "indegree": 3}, a, a) == 12


def test_syn_group_parameters(ngpu):
    # synapse groups, src/syn_model.cu:136-238: names and defaults of "stdp"
    # (src/stdp.cu:64-79) and "test_syn_model" (src/test_syn_model.cu:43-52)
    g = ngpu.CreateSynGroup("stdp")
    assert ngpu.GetSynGroupParamNames(g) == ["tau_plus", "tau_minus", "lambda", "alpha",
                                              "mu_plus", "mu_minus", "Wmax"]
    st = ngpu.GetStatus(g)
    assert st == pytest.approx({"tau_plus": 20.0, "tau_minus": 20.0, "lambda": 1.0e-4,
                                "alpha": 1.0, "mu_plus": 1.0, "mu_minus": 1.0, "Wmax": 100.0})
    ngpu.SetStatus(g, {"Wmax": 3.0, "alpha": 0.5})
    assert ngpu.GetSynGroupParam(g, "Wmax") == pytest.approx(3.0)
    assert ngpu.GetSynGroupParam(g, "alpha") == pytest.approx(0.5)
    t = ngpu.CreateSynGroup("test_syn_model", {"fact": 0.25})
    assert t.i_syn_group == g.i_syn_group + 1
    assert ngpu.GetSynGroupParamNames(t) == ["fact", "offset"]
    assert ngpu.GetSynGroupParam(t, "fact") == pytest.approx(0.25)
    with pytest.raises(ValueError, match="Unrecognized synapse parameter"):
        ngpu.GetSynGroupParam(t, "Wmax")
    # connections may name a group; an unknown group is an error
    n = ngpu.Create("aeif_cond_beta_multisynapse", 4, 1)
    ngpu.Connect(n, n, {"rule": "all_to_all"}, {"weight": 1.0, "delay": 1.0, "synapse_group": g})
    assert len(ngpu.GetConnections(n, n, syn_group=g.i_syn_group)) == 16
    assert len(ngpu.GetConnections(n, n, syn_group=0)) == 0
    with pytest.raises(ValueError, match="Unrecognized synapse group"):
        ngpu.Connect(n, n, {"rule": "one_to_one"}, {"weight": 1.0, "delay": 1.0, "synapse_group": 9})
