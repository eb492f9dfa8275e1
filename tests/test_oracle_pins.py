"""Pins of oracle functions that no reference golden vector covers (no GPU).

izhikevich_cond_beta (src/izhikevich_cond_beta_kernel.h:152-213,
src/izhikevich_cond_beta.cu:44-100) has no NEST trace in the reference
(python/test/test_izh_nest.txt is the Euler `izhikevich` model), so the
oracle's restatement is pinned by what the mathematics fixes:
  * the default state V = -70, u = b V = -14 is an equilibrium of
    0.04 V^2 + 5 V + 140 - u + I (with I = 0) and a (b V - u);
  * after one conductance input the trajectory equals an independent FP64
    integration of the same ODE (scipy's DOP853 at rtol 1e-11), started from
    the rest state with g1 = w g0, g0 from the closed-form peak of the beta
    function (computed here, not by the oracle);
  * at a spike V is reset to c and held there for round(t_ref/dt) steps while
    u relaxes as the closed form b c + (u_s - b c) exp(-a t), and u jumps by d.

The poisson_generator window (src/poiss_gen.cu:71-73: draw only when
start <= t - origin - delay <= stop, in double with float parameters) is
pinned by known answers worked out by hand below.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle

DT = 0.1


def _izh_run(status, syn=None, sim_ms=100.0, prec=0, port_status=None, rec=("V_m", "u", "g1", "g")):
    o = Oracle(prec=prec)
    n = o.create("izhikevich_cond_beta", 1, 2)
    for k, v in status.items():
        o.set(n, 1, k, [v])
    for k, vals in (port_status or {}).items():
        for p, v in enumerate(vals):
            o.set(n, 1, k, [v], port=p)
    if syn:
        sg = o.create("spike_generator", 1, 0)
        o.set_spike_times(sg, syn["times"])
        o.connect([sg], [n], [syn["port"]], [syn["weight"]], [syn["delay"]])
    recs = [o.record(n, v, 0) for v in rec]
    o.simulate(sim_ms)
    t = o.record_data(recs[0])[0].astype(np.float64)
    out = {v: o.record_data(r)[1].astype(np.float64) for v, r in zip(rec, recs)}
    return o, t, out


@pytest.mark.parametrize("prec", [0, 1])
def test_izhikevich_rest_point_is_fixed(prec):
    # 0.04*(-70)^2 + 5*(-70) + 140 = -14 = u, and b*V - u = 0
    _, t, r = _izh_run({}, sim_ms=1000.0, prec=prec)
    assert abs(r["V_m"][0] + 70.0) < 1e-6 and abs(r["u"][0] + 14.0) < 1e-5
    assert np.max(np.abs(r["V_m"] + 70.0)) < 1e-4
    assert np.max(np.abs(r["u"] + 14.0)) < 1e-4


def _beta_g0(tau_r, tau_d):
    # peak of exp(-t/tau_d) - exp(-t/tau_r) at t_p = tau_d tau_r ln(tau_d/tau_r)/(tau_d - tau_r);
    # g0 makes the conductance of a unit input peak at 1:
    # g(t) = g0 * (exp(-t/tau_d) - exp(-t/tau_r)) / (1/tau_r - 1/tau_d)
    t_p = tau_d * tau_r * math.log(tau_d / tau_r) / (tau_d - tau_r)
    return (1.0 / tau_r - 1.0 / tau_d) / (math.exp(-t_p / tau_d) - math.exp(-t_p / tau_r))


def test_beta_g0_closed_form_peak():
    # the conductance after a unit jump of g1 by g0 peaks at exactly 1
    tr, td = 2.0, 20.0
    g0 = _beta_g0(tr, td)
    t = np.linspace(0, 60, 600001)
    g = g0 * (np.exp(-t / td) - np.exp(-t / tr)) / (1.0 / tr - 1.0 / td)
    assert abs(g.max() - 1.0) < 1e-9


@pytest.mark.parametrize("prec,tol", [(1, 1e-6), (0, 5e-6)])
def test_izhikevich_subthreshold_vs_solve_ivp(prec, tol):
    from scipy.integrate import solve_ivp
    E_rev = [0.0, -85.0]
    tau_r, tau_d = [2.0, 2.0], [20.0, 20.0]
    w = 0.05
    o, t, r = _izh_run({}, syn={"times": [10.0], "port": 0, "weight": w, "delay": 1.0},
                       sim_ms=120.0, prec=prec, port_status={"E_rev": E_rev})
    # rows are recorded after the node update and before the fold
    # (src/nestgpu.cu:469-635), so the input folded at the end of step j0
    # first shows in row j0 + 1, already decayed over one step
    j = int(np.argmax(r["g1"] != 0.0))
    j0 = j - 1
    assert j0 > 0 and np.all(r["g1"][:j] == 0.0)
    a, b, I_e = 0.02, 0.2, 0.0
    g0 = _beta_g0(tau_r[0], tau_d[0])

    def rhs(_, s):
        V, u, g, g1 = s
        I_syn = g * (E_rev[0] - V)
        return [0.04 * V * V + 5.0 * V + 140.0 - u + I_syn + I_e, a * (b * V - u),
                g1 - g / tau_d[0], -g1 / tau_r[0]]

    # the rest state (an exact equilibrium, see above) with g1 = w g0 at the
    # end of step j0 (port_weight = g0, src/izhikevich_cond_beta.cu Init)
    y0 = [-70.0, -14.0, 0.0, w * g0]
    sol = solve_ivp(rhs, (t[j0], t[-1]), y0, method="DOP853", rtol=1e-11, atol=1e-12,
                    t_eval=t[j0:])
    V_ref = sol.y[0]
    assert np.ptp(V_ref) > 1.0 and V_ref.max() < 0.0   # a real, subthreshold response
    err = np.max(np.abs(r["V_m"][j0:] - V_ref) / np.abs(V_ref))
    g1_err = np.max(np.abs(r["g1"][j0 + 1:] - sol.y[3][1:]) / (w * g0))
    g_err = np.max(np.abs(r["g"][j0 + 1:] - sol.y[2][1:])) / np.max(sol.y[2])
    u_err = np.max(np.abs(r["u"][j0:] - sol.y[1]) / np.abs(sol.y[1]))
    print("izhikevich vs DOP853 max rel err: V", err, "u", u_err, "g", g_err, "g1", g1_err,
          "prec", prec)
    assert err < tol and u_err < tol and g_err < 10 * tol and g1_err < 10 * tol


def test_izhikevich_reset_and_refractory():
    t_ref = 2.0
    status = {"I_e": 10.0, "t_ref": t_ref}
    o, t, r = _izh_run(status, sim_ms=200.0, rec=("V_m", "u"))
    a, b, c, d = 0.02, 0.2, -65.0, 8.0
    ev_t, ev_n = o.events()
    assert len(ev_t) >= 3
    n_ref = int(round(t_ref / DT))
    V, u = r["V_m"], r["u"]
    for s in ev_t[:-1]:
        s = int(s)
        # V held at c for the spike's step and the n_ref - 1 steps after it
        assert np.all(V[s:s + n_ref] == np.float32(c)), (s, V[s:s + n_ref + 1])
        assert V[s + n_ref] != np.float32(c)
        # u relaxes towards b c as the closed form of du/dt = a (b c - u)
        tt = t[s:s + n_ref] - t[s]
        u_cf = b * c + (u[s] - b * c) * np.exp(-a * tt)
        assert np.max(np.abs(u[s:s + n_ref] - u_cf)) < 1e-4
        # spike-driven jump: u grew by d within the spike's step
        assert abs(u[s] - u[s - 1] - d) < 0.05


def _detector_with_poisson(rate, origin, start, stop, delay, sim_ms, prec=0):
    o = Oracle(prec=prec)
    det = o.create("spike_detector", 1, 1)
    pg = o.create("poisson_generator", 1, 0)
    for k, v in (("rate", rate), ("origin", origin), ("start", start), ("stop", stop)):
        o.set(pg, 1, k, [v])
    o.connect([pg], [det], [0], [1.0], [delay])
    rec = o.record(det, "spike_height", 0)
    o.simulate(sim_ms)
    return o.record_data(rec)[1]


def test_poisson_window_known_answer():
    # rate 2e5 Hz: lambda = 20 draws per 0.1 ms step, P(0) = 2e-9.
    # origin 5, start 10, stop 20, delay 1 (d_int 9; the edge's delay
    # 10 * 0.1f rounds to 1.0 in float): step k ends at t1 = k * 0.1f (in
    # double, k * 0.100000001490116), t_rel = t1 - 5 - 1:
    #   k = 159: t1 = 15.90000024, t_rel = 9.90000024 < 10 -> no draw
    #   k = 160: t1 = 16.00000024, t_rel = 10.00000024 >= 10 -> draw
    #   k = 259: t1 = 25.90000039, t_rel = 19.90000039 <= 20 -> draw
    #   k = 260: t1 = 26.00000039, t_rel = 20.00000039 > 20 -> no draw
    # The input of step k is folded at its end and read by the detector in
    # step k+1: record rows (row j = end of step j) 161..260 are non-zero.
    h = _detector_with_poisson(2.0e5, 5.0, 10.0, 20.0, 1.0, 40.0)
    nz = np.nonzero(h)[0]
    assert nz[0] == 161 and nz[-1] == 260 and len(nz) == 100
    assert np.all(h[nz] >= 1.0)


def test_poisson_window_defaults_and_empty():
    # defaults origin = start = 0, stop = 1e30, delay 0.1 (d_int 0): step 1
    # has t_rel = 0.1f - 0.1f = 0 >= start, so every step draws from the first
    h = _detector_with_poisson(2.0e5, 0.0, 0.0, 1.0e30, 0.1, 5.0)
    assert np.nonzero(h)[0][0] == 2 and np.all(h[2:] > 0)
    # stop before start: never
    h = _detector_with_poisson(2.0e5, 0.0, 10.0, 5.0, 0.1, 20.0)
    assert not np.any(h)
    # origin shifts the window: origin 3, start 0, stop 2, delay 0.1:
    # draws while 0 <= t1 - 3 - 0.1 <= 2, i.e. steps 31..50 (t1 = 3.1..5.0;
    # step 51 has t_rel = 2.00000009 > 2), rows 32..51
    h = _detector_with_poisson(2.0e5, 3.0, 0.0, 2.0, 0.1, 10.0)
    nz = np.nonzero(h)[0]
    assert nz[0] == 32 and nz[-1] == 51 and len(nz) == 20


def _parrot_chain(hold, backend):
    """spike_generator (t = 1.0 ms, height 2.5) -> parrot (delay 0.5, weight
    1) -> spike_detector (delay 0.1); records parrot V and detector
    spike_height."""
    if backend is None:
        o = Oracle(prec=0)
        sg = o.create("spike_generator", 1, 0)
        o.set_spike_times(sg, [1.0], [2.5])
        par = o.create("parrot_neuron", 1, 2)
        o.set(par, 1, "hold_spike_height", [hold])
        det = o.create("spike_detector", 1, 1)
        o.connect([sg], [par], [0], [1.0], [0.5])
        o.connect([par], [det], [0], [1.0], [0.1])
        rv, rs = o.record(par, "V"), o.record(det, "spike_height")
        o.simulate(4.0)
        return o.record_data(rv)[1], o.record_data(rs)[1]
    ngpu = backend
    sg = ngpu.Create("spike_generator")
    ngpu.SetStatus(sg, {"spike_times": [1.0], "spike_heights": [2.5]})
    par = ngpu.Create("parrot_neuron")
    ngpu.SetStatus(par, {"hold_spike_height": hold})
    det = ngpu.Create("spike_detector")
    ngpu.Connect(sg, par, {"rule": "one_to_one"}, {"weight": 1.0, "delay": 0.5})
    ngpu.Connect(par, det, {"rule": "one_to_one"}, {"weight": 1.0, "delay": 0.1})
    rec = ngpu.CreateRecord("", ["V", "spike_height"], [par[0], det[0]], [0, 0])
    ngpu.Simulate(4.0)
    d = np.array(ngpu.GetRecordData(rec), np.float32)
    return d[:, 1], d[:, 2]


@pytest.mark.parametrize("hold", [0.0, 1.0])
def test_parrot_neuron_known_answer(hold):
    # the generator emits in the step with time index 11 (spike time 1.0 ms,
    # src/spike_generator.cu:49-67); with d_int = 4 the input arrives in step
    # 16 and is folded at its end; the parrot reads it in step 17 (V = 2.5)
    # and re-emits (height 1, or 2.5 with hold_spike_height); d_int = 0: the
    # detector's input is folded at the end of step 18 and read in step 19
    v, h = _parrot_chain(hold, None)
    assert np.nonzero(v)[0].tolist() == [17] and v[17] == np.float32(2.5)
    assert np.nonzero(h)[0].tolist() == [19]
    assert h[19] == np.float32(2.5 if hold else 1.0)
