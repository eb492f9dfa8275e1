"""Pins of the CPU oracle (no GPU).

Every function of the oracle is checked against something it does not
restate itself: the reference's NEST golden traces (python/test/*_nest.txt
with the tolerances of the matching python/test/test_*.py), published
known-answer vectors (Philox), closed forms (linear ODE order of accuracy,
propagator P32, beta-conductance response, PSP of iaf_psc_exp), and the
reference's own known-answer delivery test (python/test/test_spike_detector.py).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from scenarios import SCENARIOS, OracleBackend, nest_trace, trace_rmse


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_nest_golden_trace(name):
    sc = SCENARIOS[name]
    _, v = OracleBackend(0).run_single(sc)
    _, vn = nest_trace(sc["golden"])
    assert len(v) == 8001
    assert trace_rmse(v, vn) < sc["tol"]


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_float_path_vs_fp64(name):
    sc = SCENARIOS[name]
    _, v32 = OracleBackend(0).run_single(sc)
    _, v64 = OracleBackend(1).run_single(sc)
    rel = np.max(np.abs(v32 - v64) / np.abs(v64))
    assert rel < 2e-6


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                    [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


@pytest.mark.parametrize("lam", [0.3, 2.0, 4.0, 23.2])
def test_poisson_sampler_statistics(lam):
    from scipy import stats
    n = 40000
    draws = np.array([O.poisson(1234, s % 97, s // 97, lam) for s in range(n)])
    assert abs(draws.mean() - lam) < 5 * math.sqrt(lam / n)
    assert abs(draws.var() - lam) < 6 * lam * math.sqrt(2.0 / n) + 0.02 * lam
    kmax = int(lam + 6 * math.sqrt(lam) + 3)
    obs = np.bincount(np.minimum(draws, kmax), minlength=kmax + 1)
    pmf = stats.poisson.pmf(np.arange(kmax + 1), lam)
    pmf[-1] = 1 - pmf[:-1].sum()
    exp = pmf * n
    keep = exp > 5
    chi2 = (((obs - exp) ** 2) / exp)[keep].sum()
    assert chi2 < stats.chi2.ppf(0.999, keep.sum() - 1)


def test_poisson_zero_rate():
    assert all(O.poisson(1, s, 7, 0.0) == 0 for s in range(100))


def test_rk5_order_linear():
    # one forced step of y' = lam*y: global-step error ~ h^6 for the 5th-order
    # solution (Cash-Karp tableau, src/rk5.cu:35-65)
    lam = -1.3
    errs = []
    hs = [0.1, 0.05, 0.025]
    for h in hs:
        y1, x, _ = O.rk5_linear(1, lam, 1.0, h)
        assert x == pytest.approx(h)
        errs.append(abs(y1 - math.exp(lam * h)))
    order = [math.log(errs[i] / errs[i + 1], 2) for i in range(2)]
    assert all(5.6 < o < 6.4 for o in order), order


def test_rk5_adaptive_lands_on_grid_and_is_accurate():
    # adaptive loop of src/rk5.h:161-174 on y' = -y: final x in [x1-h_min, x1]
    for prec in (0, 1):
        y, x, att, rej = O.rk5_linear_adaptive(prec, -1.0, 1.0, 2.0, 1e-3, 1e-4)
        assert 2.0 - 1e-4 <= x <= 2.0 + 1e-12
        # local tolerance eps=1e-6 relative to |y|+|h y'|+0.1 (src/rk5.cu:67)
        assert abs(y - math.exp(-x)) < 1e-6
        assert att >= 1


def test_propagator_32_closed_form():
    # V(h) of iaf_psc_exp for unit current: tau*tau_s/(C(tau-tau_s)) (e^-h/tau - e^-h/tau_s)
    for ts, tm, C, h in [(2.0, 10.0, 250.0, 0.1), (0.5, 10.0, 250.0, 0.1), (40.0, 10.0, 1.0, 1.0)]:
        ref = tm * ts / (C * (tm - ts)) * (math.exp(-h / tm) - math.exp(-h / ts))
        assert O.propagator_32(ts, tm, C, h) == pytest.approx(ref, rel=1e-12)
    # singular case tau == tau_s: h/C e^-h/tau
    assert O.propagator_32(10.0, 10.0, 250.0, 0.1) == pytest.approx(0.1 / 250.0 * math.exp(-0.01), rel=1e-14)
    # continuity across the switch
    a = O.propagator_32(10.0 - 1e-9, 10.0, 250.0, 0.1)
    b = O.propagator_32(10.0, 10.0, 250.0, 0.1)
    assert a == pytest.approx(b, rel=1e-7)


def _single_input_run(model, n_port, status, port, w, delay, var, prec=1, t_sim=100.0, rec_port=0):
    o = O.Oracle(prec=prec)
    n = o.create(model, 1, n_port)
    for k, v in status.items():
        if isinstance(v, list):
            for p, vv in enumerate(v):
                o.set(n, 1, k, [vv], port=p)
        else:
            o.set(n, 1, k, [v])
    sg = o.create("spike_generator", 1, 0)
    o.set_spike_times(sg, [1.0])
    o.connect([sg], [n], [port], [w], [delay])
    rec = o.record(n, var, rec_port)
    o.simulate(t_sim)
    return o.record_data(rec)


def test_beta_conductance_closed_form():
    # g(t) after one input of weight w: w*g0/(1/tr-1/td) (e^-t/td - e^-t/tr),
    # g0 chosen so the peak equals w (src/aeif_cond_beta_multisynapse.cu:72-99)
    tr, td, w = 2.0, 20.0, 3.0
    t, g = _single_input_run("aeif_cond_beta_multisynapse", 1,
                             {"tau_rise": [tr], "tau_decay": [td]}, 0, w, 1.0, "g")
    # the generator pushes its 1.0 ms spike in loop step it=10, whose time
    # index is 11 (src/spike_generator.cu:49-67, src/nestgpu.cu:486-490);
    # delay 1 ms (d_int 9) -> folded at the end of time index 11+1+9 = 21
    t0 = 2.1
    tt = t.astype(np.float64)
    tp = td * tr * math.log(td / tr) / (td - tr)
    g0 = (1 / tr - 1 / td) / (math.exp(-tp / td) - math.exp(-tp / tr))
    ana = np.where(tt > t0 + 1e-6, w * g0 / (1 / tr - 1 / td) *
                   (np.exp(-(tt - t0) / td) - np.exp(-(tt - t0) / tr)), 0.0)
    assert np.max(np.abs(g - ana)) < 1e-5 * w
    assert abs(np.max(ana) - w) < 1e-3 * w


def test_iaf_psc_exp_psp_closed_form():
    # exact integration: V(t) = I0*tau_m*tau_s/(C(tau_m - tau_s))(e^-t/tau_m - e^-t/tau_s)
    I0, tm, ts, C = 500.0, 10.0, 2.0, 250.0
    t, v = _single_input_run("iaf_psc_exp", 2, {}, 0, I0, 1.0, "V_m_rel", prec=0)
    t0 = 2.1
    tt = t.astype(np.float64)
    ana = np.where(tt > t0 + 1e-6, I0 * tm * ts / (C * (tm - ts)) *
                   (np.exp(-(tt - t0) / tm) - np.exp(-(tt - t0) / ts)), 0.0)
    assert np.max(np.abs(v - ana)) < 1e-5 * np.max(np.abs(ana))


def test_spike_detector_known_answer():
    # python/test/test_spike_detector.py:5-51: detector rows at +101/201/301
    o = O.Oracle(prec=0)
    n = o.create("aeif_cond_beta_multisynapse", 3, 1)
    o.set(n, 3, "I_e", [1000.0, 1000.0, 1000.0])
    sd = o.create("spike_detector", 1, 1)
    for k, (w, d) in enumerate([(1.0, 10.0), (2.0, 20.0), (3.0, 30.0)]):
        o.connect([n + k], [sd], [0], [w], [d])
    rn = o.record(n, "spike", 0)
    rs = o.record(sd, "spike_height", 0)
    o.simulate(1000.0)
    _, sn = o.record_data(rn)
    _, ss = o.record_data(rs)
    hits = 0
    for i in range(len(sn) - 400):
        if sn[i] > 0.5:
            hits += 1
            assert abs(ss[i + 101] - 1.0) < 1e-6
            assert abs(ss[i + 201] - 2.0) < 1e-6
            assert abs(ss[i + 301] - 3.0) < 1e-6
    assert hits > 5


def test_routing_brute_force():
    # every (spike, synapse) arrives exactly delay after emission, at the
    # synapse's port, with weight*height (src/spike_buffer.cu:247-317,
    # src/get_spike.cu:60-97): a detector per (source, delay) pair reads it.
    o = O.Oracle(prec=0)
    sg = o.create("spike_generator", 2, 0)
    det = o.create("spike_detector", 6, 1)
    times = [[1.0, 2.5, 7.0], [3.0, 3.1]]
    for k in range(2):
        o.set_spike_times(sg + k, times[k], [1.0 + k, 0.5 + k])
    o.set(det, 6, "hold_spike_height", [1.0] * 6)
    conns = []
    for k in range(2):
        for j, d in enumerate([0.1, 0.4, 1.3]):
            tgt = det + 3 * k + j
            w = 0.25 * (j + 1)
            conns.append((sg + k, tgt, w, d))
            o.connect([sg + k], [tgt], [0], [w], [d])
    recs = [o.record(det + i, "spike_height", 0) for i in range(6)]
    o.simulate(12.0)
    for (s, tgt, w, d) in conns:
        k = s - sg
        _, v = o.record_data(recs[tgt - det])
        expect = np.zeros_like(v)
        for ts, hgt in zip(times[k], [1.0 + k, 0.5 + k]):
            emit_row = int(round(ts / 0.1))          # generator pushes at step it
            d_int = int(round(d / 0.1)) - 1
            # folded at end of step it+1+d_int, detector sees it one row later
            row = emit_row + 1 + d_int + 1 + 1
            expect[row] += w * hgt
        np.testing.assert_array_equal(v, expect.astype(np.float32))


def test_step_factors_are_correctly_rounded_powers():
    # err^-0.25 and err^-0.2 of the controller (src/rk5.cu:68-70) rounded from
    # the exact power; compare with mpmath-free high precision: double pow,
    # whose error (< 1 ulp of double) cannot move the float rounding except
    # at ties, which the random sample does not hit
    rng = np.random.default_rng(7)
    xs = np.concatenate([10.0 ** rng.uniform(-4, 6, 20000), [1.0, 1.889568e-4, 2.0, 1e6, 3.5]])
    xs = xs.astype(np.float32)
    for x in xs:
        d = np.float32(np.float64(x) ** -0.25)
        i = np.float32(np.float64(x) ** -0.2)
        assert O.step_factor_dec(float(x)) == d
        assert O.step_factor_inc(float(x)) == i


def test_step_factors_strided_sweep_of_all_positive_floats():
    # every 127th positive float bit pattern (denormals to FLT_MAX, 17M values)
    # against the double power rounded to float; a tie closer than double
    # rounding is the only way the two may differ, and none is expected
    u = np.arange(1, 0x7F800000, 127, dtype=np.uint32)
    x = u.view(np.float32)
    dec, inc = O.step_factors(x)
    xd = x.astype(np.float64)
    np.testing.assert_array_equal(dec, (xd ** -0.25).astype(np.float32))
    np.testing.assert_array_equal(inc, (xd ** -0.2).astype(np.float32))
    d0, i0 = O.step_factors(np.array([0.0, np.inf, np.nan], np.float32))
    assert d0[0] == np.inf and i0[0] == np.inf and d0[1] == 0 and i0[1] == 0
    assert np.isnan(d0[2]) and np.isnan(i0[2])


def test_exp_model_within_one_ulp_of_exp():
    # the AdEx exponential term (src/aeif_cond_beta_multisynapse_kernel.h:177)
    # in float (Cody-Waite + degree-7 polynomial): every float in a strided
    # sweep of [-110, 100] within 1 ulp of exp evaluated in float64, almost
    # all correctly rounded (a dropped or wrong polynomial term breaks both);
    # overflow, underflow and the special values of the IEEE exp
    u = np.arange(0, 0x80000000, 1009, dtype=np.uint64).astype(np.uint32)
    x = np.concatenate([u.view(np.float32), (u | np.uint32(0x80000000)).view(np.float32)])
    x = x[np.isfinite(x) & (x > -110) & (x < 100)]
    got = O.exp_model(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    fin = ref < np.float64(np.finfo(np.float32).max)
    # ulp of the float result (subnormal spacing below the normal range)
    r32 = ref[fin].astype(np.float32)
    ulp = np.spacing(np.maximum(np.abs(r32), np.float32(np.finfo(np.float32).tiny))).astype(np.float64)
    err = np.abs(got[fin] - ref[fin]) / ulp
    print("exp: %d points, max error %.3f ulp, correctly rounded %.4f"
          % (err.size, err.max(), np.mean(got[fin] == r32)))
    assert err.max() < 1.0
    assert np.mean(got[fin] == r32) > 0.99
    assert np.all(np.isinf(got[~fin]))
    assert O.exp_model(0.0) == 1.0 and O.exp_model(-0.0) == 1.0
    assert O.exp_model(1.0) == np.float32(np.e)
    assert O.exp_model(88.73) == float("inf") and O.exp_model(100.0) == float("inf")
    assert O.exp_model(-104.0) == 0.0 and O.exp_model(-200.0) == 0.0
    assert 0.0 < O.exp_model(-103.0) < np.finfo(np.float32).tiny   # subnormal result
    assert np.isnan(O.exp_model(float("nan")))
