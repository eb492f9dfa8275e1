"""Device-side connectivity.

* The random connection rules drawn on the GPU (fixed_indegree and
  fixed_outdegree with both of their samplers, fixed_total_number) give the
  same connections as the host loops: the same Philox words, taken in the
  same order (src/connect_rules.h:161-338 are the reference's rules).
* The connection store flattened on the GPU (a stable radix sort by source
  and delay slot) gives the synapse table of the host counting sort entry
  for entry -- the order the reference's flattening produces
  (src/connect.cu:56-74, src/spike_buffer.cu:399-437): by source, by delay
  inside a source, creation order inside a (source, delay) group."""
import ctypes
import os

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


def _connect_arrays(ngpu, rule, src, tgt, n_conn, port, rng, **spec):
    w = rng.uniform(0.5, 1.5, n_conn).astype(np.float32)
    d = (rng.integers(1, 31, n_conn) * 0.1).astype(np.float32)
    lib = ngpu._nestgpu
    ngpu.ConnSpecInit()
    ngpu.SynSpecInit()
    ngpu.SetConnSpecParam("rule", rule)
    for k, v in spec.items():
        ngpu.SetConnSpecParam(k, v)
    ngpu.SetSynSpecIntParam("receptor", port)
    lib.NESTGPU_SetSynSpecFloatPtParam(b"weight_array", w.ctypes.data_as(ctypes.c_void_p))
    lib.NESTGPU_SetSynSpecFloatPtParam(b"delay_array", d.ctypes.data_as(ctypes.c_void_p))
    lib.NESTGPU_ConnectSeqSeq(src.i0, src.n, tgt.i0, tgt.n)
    ngpu._check()


def _build(ngpu, seed):
    ngpu.ResetKernel()
    ngpu.SetVerbosityLevel(0)
    ngpu.SetRandomSeed(seed)
    a = ngpu.Create("iaf_psc_exp", 2000, 2)
    b = ngpu.Create("iaf_psc_exp", 1500, 2)
    rng = np.random.default_rng(seed)
    # rule codes: 0 one_to_one, 1 all_to_all, 2 fixed_total_number,
    # 3 fixed_indegree, 4 fixed_outdegree
    _connect_arrays(ngpu, 3, a, b, 1500 * 300, 0, rng, indegree=300)
    _connect_arrays(ngpu, 4, b, a, 1500 * 200, 1, rng, outdegree=200)
    _connect_arrays(ngpu, 2, a, a, 250000, 0, rng, total_num=250000)
    _connect_arrays(ngpu, 1, ngpu.NodeSeq(b.i0, 100), ngpu.NodeSeq(a.i0, 900), 90000, 1, rng)
    _connect_arrays(ngpu, 0, ngpu.NodeSeq(a.i0, 1500), b, 1500, 0, rng)
    # pool smaller than 5 x degree: the partial Fisher-Yates sampler
    _connect_arrays(ngpu, 3, ngpu.NodeSeq(b.i0, 1000), ngpu.NodeSeq(a.i0, 500), 500 * 400, 1, rng,
                    indegree=400)
    _connect_arrays(ngpu, 4, ngpu.NodeSeq(a.i0, 300), ngpu.NodeSeq(b.i0, 600), 300 * 250, 0, rng,
                    outdegree=250)
    # single connections interleaved with the rules
    ngpu.Connect([a[0]], [b[0]], {"rule": "one_to_one"}, {"weight": 2.0, "delay": 0.5, "receptor": 1})
    n = a.n + b.n
    inc = ngpu.GetIncoming(list(range(n)))
    return n, inc


@pytest.mark.parametrize("seed", [11, 12])
def test_device_rules_match_host_rules(ngpu, monkeypatch, seed):
    res = []
    monkeypatch.setenv("NGPU_DEVICE_SORT", "0")
    for mode in ("0", "1"):
        monkeypatch.setenv("NGPU_DEVICE_CONNECT", mode)
        res.append(_build(ngpu, seed))
    (n0, h), (n1, dv) = res
    assert n0 == n1
    for key in ("source", "target", "port", "weight", "delay", "index"):
        np.testing.assert_array_equal(h[key], dv[key], err_msg=key)


@pytest.mark.parametrize("seed", [11, 12])
def test_device_sort_matches_host_sort(ngpu, monkeypatch, seed):
    res = []
    monkeypatch.setenv("NGPU_DEVICE_CONNECT", "1")
    for mode in ("0", "1"):
        monkeypatch.setenv("NGPU_DEVICE_SORT", mode)
        res.append(_build(ngpu, seed))
    (n0, h), (n1, dv) = res
    assert n0 == n1
    m = len(h["source"])
    assert m == 1500 * 300 + 1500 * 200 + 250000 + 90000 + 1500 + 500 * 400 + 300 * 250 + 1
    for key in ("source", "target", "port", "weight", "delay", "index"):
        np.testing.assert_array_equal(h[key], dv[key], err_msg=key)
    # the table's order: source ascending, delay ascending inside a source
    s, d = dv["source"], dv["delay"]
    order = np.lexsort((d, s))
    assert np.array_equal(s[order], s) and np.array_equal(d[order], d)


def test_device_sort_default_from_a_million_connections(ngpu, monkeypatch):
    # without the switch the device sort runs from 2^20 connections on; a
    # simulation on the device-sorted table runs and delivers
    monkeypatch.delenv("NGPU_DEVICE_SORT", raising=False)
    monkeypatch.delenv("NGPU_DEVICE_CONNECT", raising=False)
    n, inc = _build(ngpu, 5)
    assert len(inc["source"]) > (1 << 20)
    pg = ngpu.Create("poisson_generator")
    ngpu.SetStatus(pg, "rate", 20000.0)
    ngpu.Connect(pg, ngpu.NodeSeq(0, n), {"rule": "all_to_all"}, {"weight": 300.0, "delay": 0.1})
    ngpu.Simulate(20.0)
    st = ngpu.GetStepStats()
    assert st["spikes"] > 0 and st["syn_events"] > 0
