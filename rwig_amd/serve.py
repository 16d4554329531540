"""Python handle on the serving engine of librwig_hip.so (rwig_serve_*).

A session keeps cost matrices in HBM and solves Sinkhorn and barycenter
requests as they arrive, batching those that share a (cost, reg) pair into
one MFMA GEMM per Gibbs product. Requests are admitted in arrival order while
the session's HBM budget holds them.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3

STAT_NAMES = ("completed", "failed", "queued", "running", "p50_ms", "p99_ms", "mean_ms",
              "solves_per_s", "bytes_in_use", "peak_bytes", "budget_bytes", "cycles", "groups",
              "mean_queue_ms", "max_ms")


def _p(x):
    return x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(x):
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


def footprint(M, N, columns=32):
    """(Gibbs kernel bytes, group bytes at `columns` columns) that a session
    charges against its HBM budget for one (cost, reg) pair."""
    kb, gb = ctypes.c_double(), ctypes.c_double()
    check(lib().rwig_serve_footprint(int(M), int(N), int(columns), ctypes.byref(kb),
                                     ctypes.byref(gb)))
    return kb.value, gb.value


class ServeSession:
    """rwig_serve_create / _destroy; use as a context manager."""

    def __init__(self, hbm_budget_bytes=8 << 30, max_group_columns=512, absorb_every=20,
                 switch_thr=1e-2, graphs=True):
        h = ctypes.c_void_p()
        check(lib().rwig_serve_create(float(hbm_budget_bytes), int(max_group_columns),
                                      int(absorb_every), float(switch_thr),
                                      0 if graphs else 1, ctypes.byref(h)))
        self._h = h
        self._shapes = {}
        self._costs = {}
        self._kinds = {}

    def close(self):
        if self._h:
            check(lib().rwig_serve_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def register_cost(self, C):
        """Copies C (M x N) to the device; returns its id."""
        C = _f64(C)
        cid = ctypes.c_int()
        check(lib().rwig_serve_register_cost(self._h, _p(C), C.shape[0], C.shape[1],
                                             ctypes.byref(cid)))
        self._shapes[cid.value] = C.shape
        return cid.value

    def submit_sinkhorn(self, cost_id, a, b, reg, max_iter=1000, zero_tol=1e-6):
        a = _f64(a).ravel()
        b = _f64(b).ravel()
        rid = ctypes.c_long()
        check(lib().rwig_serve_submit_sinkhorn(self._h, int(cost_id), float(reg), _p(a), a.size,
                                               _p(b), b.size, int(max_iter), float(zero_tol),
                                               ctypes.byref(rid)))
        self._costs[rid.value] = cost_id
        self._kinds[rid.value] = ("sinkhorn", 1)
        return rid.value

    def submit_barycenter(self, cost_id, A, w, reg, max_iter=1000, zero_tol=1e-6):
        A = _f64(A)
        w = _f64(w).ravel()
        rid = ctypes.c_long()
        check(lib().rwig_serve_submit_barycenter(self._h, int(cost_id), float(reg), _p(A),
                                                 A.shape[0], A.shape[1], _p(w), int(max_iter),
                                                 float(zero_tol), ctypes.byref(rid)))
        self._costs[rid.value] = cost_id
        self._kinds[rid.value] = ("barycenter", A.shape[1])
        return rid.value

    def wait(self, request_id, timeout_s=-1.0):
        st = ctypes.c_int()
        check(lib().rwig_serve_wait(self._h, int(request_id), float(timeout_s),
                                    ctypes.byref(st)))
        return st.value

    def result(self, request_id):
        """Waits for the request and returns its result dict (then forgotten).

        return_status is the reference's (src/sinkhorn.cpp:88-95): 0 converged,
        1 max_iter reached, 2 the solve broke down (non-finite error)."""
        self.wait(request_id)
        M, N = self._shapes[self._costs.pop(request_id)]
        kind, S = self._kinds.pop(request_id)
        x = np.zeros(M if kind == "sinkhorn" else N)
        y = np.zeros(N if kind == "sinkhorn" else (M + N) * S)
        loss, err, lat = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        it, status = ctypes.c_int(), ctypes.c_int()
        check(lib().rwig_serve_result(self._h, int(request_id), _p(x), _p(y), ctypes.byref(loss),
                                      ctypes.byref(it), ctypes.byref(err), ctypes.byref(lat),
                                      ctypes.byref(status)))
        out = {"iter": it.value, "err": err.value, "latency_ms": lat.value,
               "return_status": status.value}
        if kind == "sinkhorn":
            out.update(f=x, g=y, loss=loss.value)
        else:
            out["b"] = x
            out["F"] = y[:M * S].reshape((M, S), order="F")
            out["G"] = y[M * S:].reshape((N, S), order="F")
        return out

    def stats(self):
        out = np.zeros(len(STAT_NAMES))
        check(lib().rwig_serve_stats(self._h, _p(out), out.size))
        return dict(zip(STAT_NAMES, out.tolist()))

    def reset_stats(self):
        check(lib().rwig_serve_reset_stats(self._h))
