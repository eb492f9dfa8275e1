"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (liboracle.so).

The oracle is a host C++ restatement of the reference NEST GPU step (see
oracle.cpp for the file:line map).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, as the checker; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, ip, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
        L.orc_new.restype = vp
        L.orc_new.argtypes = [ctypes.c_float, ctypes.c_ulonglong, ctypes.c_int]
        L.orc_free.argtypes = [vp]
        L.orc_error.restype = ctypes.c_char_p
        L.orc_error.argtypes = [vp]
        L.orc_create.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        L.orc_set.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, fp]
        L.orc_get.restype = ctypes.c_double
        L.orc_get.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.orc_set_spike_times.argtypes = [vp, ctypes.c_int, ctypes.c_int, fp, fp]
        L.orc_connect.argtypes = [vp, ctypes.c_longlong, ip, ip,
                                  ctypes.POINTER(ctypes.c_ubyte), fp, fp]
        L.orc_connect_stream.argtypes = [vp, ctypes.c_longlong, ip, ip,
                                         ctypes.POINTER(ctypes.c_ubyte), fp, fp, vp]
        L.orc_record.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.orc_simulate.argtypes = [vp, ctypes.c_float]
        L.orc_record_rows.restype = ctypes.c_longlong
        L.orc_record_rows.argtypes = [vp, ctypes.c_int]
        L.orc_record_copy.argtypes = [vp, ctypes.c_int, fp, fp]
        L.orc_n_events.restype = ctypes.c_longlong
        L.orc_n_events.argtypes = [vp]
        L.orc_events.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ip]
        L.orc_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
        L.orc_rk5_linear.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.orc_rk5_linear_adaptive.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_double)]
        L.orc_philox.argtypes = [ctypes.POINTER(ctypes.c_uint), ctypes.c_uint, ctypes.c_uint,
                                 ctypes.POINTER(ctypes.c_uint)]
        L.orc_poisson.restype = ctypes.c_int
        L.orc_poisson.argtypes = [ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_longlong, ctypes.c_double]
        L.orc_propagator_32.restype = ctypes.c_double
        L.orc_propagator_32.argtypes = [ctypes.c_double] * 4
        L.orc_threads.restype = ctypes.c_int
        L.orc_step_factor_dec.restype = ctypes.c_float
        L.orc_step_factor_dec.argtypes = [ctypes.c_float]
        L.orc_step_factor_inc.restype = ctypes.c_float
        L.orc_step_factor_inc.argtypes = [ctypes.c_float]
        L.orc_exp_model.restype = ctypes.c_float
        L.orc_exp_model.argtypes = [ctypes.c_float]
        L.orc_exp_model_n.restype = None
        L.orc_exp_model_n.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                      ctypes.c_longlong]
        L.orc_mlp_n_params.restype = ctypes.c_int
        L.orc_mlp_forward.restype = None
        L.orc_mlp_forward.argtypes = [vp, vp, ctypes.c_longlong, ctypes.c_int, vp]
        L.orc_mlp_loss_grad.restype = None
        L.orc_mlp_loss_grad.argtypes = [vp, vp, vp, vp, ctypes.c_longlong, ctypes.c_int, vp, vp]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class Oracle:
    """One oracle network.  prec=0: reference float path, prec=1: FP64."""

    def __init__(self, dt=0.1, seed=54321 + 12345, prec=0):
        self.L = lib()
        self.h = self.L.orc_new(dt, seed, prec)

    def __del__(self):
        try:
            self.L.orc_free(self.h)
        except Exception:
            pass

    def _check(self, ret):
        if ret < 0:
            raise ValueError(self.L.orc_error(self.h).decode())
        return ret

    def create(self, model, n=1, n_port=1):
        return self._check(self.L.orc_create(self.h, model.encode(), n, n_port))

    def set(self, node0, n, name, vals, port=0):
        v = _f32(np.broadcast_to(np.asarray(vals, dtype=np.float32), (n,)))
        self._check(self.L.orc_set(self.h, node0, n, name.encode(), port,
                                   v.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def get(self, node, name, port=0):
        return self.L.orc_get(self.h, node, name.encode(), port)

    def set_spike_times(self, node, times, heights=None):
        t = _f32(times)
        hp = None
        if heights is not None:
            hh = _f32(heights)
            hp = hh.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self._check(self.L.orc_set_spike_times(self.h, node, len(t),
                                               t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), hp))

    def connect(self, src, tgt, port, weight, delay, stream=None):
        """Connections src[k] -> tgt[k]; stream (optional): explicit Poisson
        stream ids for the connections of poisson generators."""
        src = np.ascontiguousarray(src, dtype=np.int32)
        n = len(src)
        tgt = np.ascontiguousarray(np.broadcast_to(tgt, (n,)), dtype=np.int32)
        port = np.ascontiguousarray(np.broadcast_to(port, (n,)), dtype=np.uint8)
        w = _f32(np.broadcast_to(np.asarray(weight, dtype=np.float32), (n,)))
        d = _f32(np.broadcast_to(np.asarray(delay, dtype=np.float32), (n,)))
        args = [self.h, n, src.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                tgt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                port.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                d.ctypes.data_as(ctypes.POINTER(ctypes.c_float))]
        if stream is None:
            self._check(self.L.orc_connect(*args))
        else:
            st = np.ascontiguousarray(np.broadcast_to(stream, (n,)), dtype=np.int64)
            self._check(self.L.orc_connect_stream(*args, st.ctypes.data_as(ctypes.c_void_p)))

    def record(self, node, var, port=0):
        return self._check(self.L.orc_record(self.h, node, var.encode(), port))

    def simulate(self, t):
        self._check(self.L.orc_simulate(self.h, t))

    def record_data(self, rec):
        n = self.L.orc_record_rows(self.h, rec)
        t = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        self.L.orc_record_copy(self.h, rec, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return t, v

    def events(self):
        """All emitted spikes as (time_index, node), sorted."""
        n = self.L.orc_n_events(self.h)
        t = np.zeros(n, np.int64)
        nd = np.zeros(n, np.int32)
        self.L.orc_events(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                          nd.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        o = np.lexsort((nd, t))
        return t[o], nd[o]

    def stats(self):
        s = (ctypes.c_longlong * 4)()
        self.L.orc_stats(self.h, s)
        return {"attempts": s[0], "rejected": s[1], "substeps": s[2], "syn_events": s[3]}


def rk5_linear(prec, lam, y0, h):
    out = (ctypes.c_double * 3)()
    lib().orc_rk5_linear(prec, lam, y0, h, out)
    return out[0], out[1], out[2]


def rk5_linear_adaptive(prec, lam, y0, x1, h0, h_min):
    out = (ctypes.c_double * 4)()
    lib().orc_rk5_linear_adaptive(prec, lam, y0, x1, h0, h_min, out)
    return out[0], out[1], int(out[2]), int(out[3])


def philox(ctr, key):
    c = (ctypes.c_uint * 4)(*ctr)
    o = (ctypes.c_uint * 4)()
    lib().orc_philox(c, key[0], key[1], o)
    return list(o)


def poisson(seed, stream, t_idx, lam):
    return lib().orc_poisson(seed, stream, t_idx, lam)


def propagator_32(tau_syn, tau, C, h):
    return lib().orc_propagator_32(tau_syn, tau, C, h)


def threads():
    return lib().orc_threads()


def step_factor_dec(e):
    return lib().orc_step_factor_dec(e)


def step_factor_inc(e):
    return lib().orc_step_factor_inc(e)


def step_factors(x):
    """Vectorised step_factor_dec / step_factor_inc over a float32 array."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    dec = np.empty_like(x)
    inc = np.empty_like(x)
    f = lib().orc_step_factors
    f.restype = None
    f(x.ctypes.data_as(ctypes.c_void_p), ctypes.c_longlong(x.size),
      dec.ctypes.data_as(ctypes.c_void_p), inc.ctypes.data_as(ctypes.c_void_p))
    return dec, inc


def mlp_n_params():
    """Length of the step-size predictor's parameter vector."""
    return lib().orc_mlp_n_params()


def mlp_forward(params, raw, quant=1):
    """Predictor output log2(h/dt) for raw feature rows [n, 8] (FP64 host)."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    X = np.ascontiguousarray(raw, dtype=np.float32).reshape(-1, 8)
    out = np.empty(X.shape[0], np.float64)
    lib().orc_mlp_forward(P.ctypes.data, X.ctypes.data, X.shape[0], quant, out.ctypes.data)
    return out


def mlp_loss_grad(params, raw, target, cens, quant=1):
    """Mean censored squared loss and its gradient (FP64 host)."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    X = np.ascontiguousarray(raw, dtype=np.float32).reshape(-1, 8)
    T = np.ascontiguousarray(target, dtype=np.float32)
    C = np.ascontiguousarray(cens, dtype=np.uint8)
    g = np.empty(P.size, np.float64)
    loss = ctypes.c_double()
    lib().orc_mlp_loss_grad(P.ctypes.data, X.ctypes.data, T.ctypes.data, C.ctypes.data,
                            X.shape[0], quant, ctypes.addressof(loss), g.ctypes.data)
    return loss.value, g


def exp_model(x):
    """Float-path exponential of the AdEx models: a float for a scalar, an
    array for an array."""
    if np.isscalar(x):
        return lib().orc_exp_model(x)
    a = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(a)
    lib().orc_exp_model_n(a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size)
    return out
