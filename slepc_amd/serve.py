"""Python front-end of the serving loop of libslepc_amd.so (SlepcServer* in
include/slepc_amd.h): eigenproblems arrive over time, are admitted against an
HBM budget and solved concurrently, each by the unchanged Krylov-Schur solver
on its own stream.  The loop itself runs in C; this module only submits
requests at their arrival times, calls SlepcServerProgress and collects
results."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import eps as _eps
from . import matgen

_P = C.POINTER


class Request(C.Structure):
    _fields_ = [("n", C.c_int), ("rowptr", C.c_void_p), ("colidx", C.c_void_p), ("val", C.c_void_p),
                ("val32", C.c_void_p),
                ("nev", C.c_int), ("ncv", C.c_int), ("max_it", C.c_int), ("tol", C.c_double),
                ("which", C.c_int), ("hermitian", C.c_int), ("v0", C.c_void_p), ("want_vectors", C.c_int),
                ("mixed_tol", C.c_double), ("t_arrival", C.c_double), ("twosided", C.c_int)]


class Result(C.Structure):
    _fields_ = [("nconv", C.c_int), ("nret", C.c_int), ("ngood", C.c_int), ("its", C.c_int),
                ("reason", C.c_int), ("nmatmult", C.c_int64), ("gs_passes", C.c_int64),
                ("t_arrival", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("t_host_ds", C.c_double), ("error", C.c_int)]


def clock() -> float:
    L = _eps.lib()
    L.SlepcServerClock.restype = C.c_double
    return L.SlepcServerClock()


class Server:
    """SlepcServer: ``budget`` bytes of HBM for resident requests, at most
    ``max_active`` solves on the device at once."""

    def __init__(self, budget: int, max_active: int = 8, grouped: bool = True):
        _eps.initialize()
        self.h = C.c_void_p()
        _eps._chk(_eps.lib().SlepcServerCreate(C.c_size_t(int(budget)), C.c_int(max_active), C.byref(self.h)),
                  "SlepcServerCreate")
        _eps._chk(_eps.lib().SlepcServerSetGrouped(self.h, C.c_int(int(grouped))), "SlepcServerSetGrouped")
        self._keep = {}
        self._meta = {}

    def submit(self, csr, nev: int, ncv: int, tol: float = 1e-10, which: int = 1, hermitian: bool = True,
               v0=None, want_vectors: bool = False, t_arrival: float = 0.0, max_it: int = 0,
               mixed_tol: float = 0.0, twosided: bool = False) -> int:
        """FP32 values in csr[2] keep the operator in FP32; mixed_tol > tol adds the
        FP32-basis first phase; twosided runs two-sided Krylov-Schur (FP64 values)."""
        rowptr = np.ascontiguousarray(csr[0], dtype=np.int32)
        colidx = np.ascontiguousarray(csr[1], dtype=np.int32)
        f32 = np.asarray(csr[2]).dtype == np.float32
        val = np.ascontiguousarray(csr[2], dtype=np.float32 if f32 else np.float64)
        n = len(rowptr) - 1
        v0a = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
        r = Request(n=n, rowptr=rowptr.ctypes.data, colidx=colidx.ctypes.data,
                    val=None if f32 else val.ctypes.data, val32=val.ctypes.data if f32 else None, nev=nev,
                    ncv=ncv, max_it=max_it, tol=tol, which=which, hermitian=int(hermitian),
                    v0=None if v0a is None else v0a.ctypes.data, want_vectors=int(want_vectors),
                    mixed_tol=mixed_tol, t_arrival=t_arrival, twosided=int(twosided))
        rid = C.c_int()
        _eps._chk(_eps.lib().SlepcServerSubmit(self.h, C.byref(r), C.byref(rid)), "SlepcServerSubmit")
        self._keep[rid.value] = (rowptr, colidx, val, v0a)   # the C side reads them until the solve ends
        self._meta[rid.value] = (n, nev, want_vectors)
        return rid.value

    def progress(self):
        moved, na, nq, nf = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _eps._chk(_eps.lib().SlepcServerProgress(self.h, C.byref(moved), C.byref(na), C.byref(nq), C.byref(nf)),
                  "SlepcServerProgress")
        return bool(moved.value), na.value, nq.value, nf.value

    def result(self, rid: int):
        n, nev, want = self._meta[rid]
        done = C.c_int()
        res = Result()
        eigr = np.zeros(nev)
        eigi = np.zeros(nev)
        err = np.zeros(nev)
        X = np.zeros((nev, n)) if want else None
        dp = lambda a: None if a is None else a.ctypes.data_as(_P(C.c_double))
        _eps._chk(_eps.lib().SlepcServerGetResult(self.h, C.c_int(rid), C.byref(done), C.byref(res), dp(eigr),
                                                  dp(eigi), dp(err), dp(X)), "SlepcServerGetResult")
        if not done.value:
            return None
        self._keep.pop(rid, None)
        k = res.nret
        out = dict(nconv=res.nconv, nret=k, ngood=res.ngood, its=res.its, reason=res.reason,
                   nmatmult=res.nmatmult, gs_passes=res.gs_passes, t_arrival=res.t_arrival,
                   t_admit=res.t_admit, t_done=res.t_done, latency=res.t_done - res.t_arrival, t_host_ds=res.t_host_ds,
                   error=res.error, eigr=eigr[:k], eigi=eigi[:k], err=err[:k])
        if want:
            out["X"] = X[:k].T.copy()
        return out

    def stats(self) -> dict:
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _eps._chk(_eps.lib().SlepcServerGetStats(self.h, C.byref(a), C.byref(b), C.byref(c)), "SlepcServerGetStats")
        gs, gm = C.c_int64(), C.c_int64()
        _eps._chk(_eps.lib().SlepcServerGetGroupStats(self.h, C.byref(gs), C.byref(gm)), "SlepcServerGetGroupStats")
        t = (C.c_double * 5)()
        _eps._chk(_eps.lib().SlepcServerGetHostTimes(self.h, t), "SlepcServerGetHostTimes")
        return dict(peak_reserved=a.value, pool_peak=b.value, budget=c.value, group_steps=gs.value,
                    group_member_steps=gm.value,
                    host_s=dict(admit=t[0], grouped_steps=t[1], restarts=t[2], retire=t[3], poll=t[4]))

    def close(self):
        if self.h:
            _eps._chk(_eps.lib().SlepcServerDestroy(C.byref(self.h)), "SlepcServerDestroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_stream(server: Server, requests, make_csr, tol: float = 1e-10, want_vectors: bool = False,
               time_scale: float = 1.0, idle_sleep: float = 0.0):
    """Feeds ``requests`` (dicts with n, nev, ncv, t_arrival, v0_seed, ...) to the
    server at their arrival times (offsets in seconds from now, times
    ``time_scale``) and runs the loop until every request has finished.
    ``make_csr(req)`` builds a request's matrix; it is called ahead of the
    arrival time, outside the latency.  Returns the results in request order
    plus the wall time of the run."""
    mats = [None] * len(requests)
    nxt = 0
    ids = [None] * len(requests)
    # matrices and start vectors are prepared before the clock starts
    for i, r in enumerate(requests):
        mats[i] = (make_csr(r), matgen.uniform_vector(r["v0_seed"], r["n"]))
    t0 = clock()
    remaining = len(requests)
    results = [None] * len(requests)
    pending = set()
    seen_finished = 0
    while remaining:
        now = clock()
        while nxt < len(requests) and t0 + time_scale * requests[nxt]["t_arrival"] <= now:
            r = requests[nxt]
            csr, v0 = mats[nxt]
            ids[nxt] = server.submit(csr, r["nev"], r["ncv"], tol=tol, v0=v0, want_vectors=want_vectors,
                                     t_arrival=t0 + time_scale * r["t_arrival"], which=r.get("which", 1),
                                     hermitian=r.get("hermitian", True), mixed_tol=r.get("mixed_tol", 0.0),
                                     twosided=r.get("twosided", False))
            pending.add(nxt)
            mats[nxt] = None
            nxt += 1
        moved, na, nq, nf = server.progress()
        if nf > seen_finished:   # collect only when something finished
            seen_finished = nf
            for i in list(pending):
                res = server.result(ids[i])
                if res is not None:
                    results[i] = res
                    pending.discard(i)
                    remaining -= 1
        if not moved and idle_sleep > 0:
            time.sleep(idle_sleep)
    return results, clock() - t0


def run_stream_native(server: Server, requests, make_csr, tol: float = 1e-10, want_vectors: bool = False,
                      time_scale: float = 1.0):
    """The same stream as run_stream, driven by SlepcServerRun: submission at the
    arrival times, the progress loop and result collection all run in C, so no
    Python work sits between the server's steps.  Returns the results in request
    order and the wall time."""
    n = len(requests)
    keep, outs = [], []
    reqs = (Request * max(n, 1))()
    res = (Result * max(n, 1))()
    P = C.POINTER(C.c_double)
    er, ei, ee = (P * max(n, 1))(), (P * max(n, 1))(), (P * max(n, 1))()
    xs = (P * max(n, 1))() if want_vectors else None
    for i, r in enumerate(requests):
        csr = make_csr(r)
        rowptr = np.ascontiguousarray(csr[0], dtype=np.int32)
        colidx = np.ascontiguousarray(csr[1], dtype=np.int32)
        f32 = np.asarray(csr[2]).dtype == np.float32
        val = np.ascontiguousarray(csr[2], dtype=np.float32 if f32 else np.float64)
        v0 = np.ascontiguousarray(matgen.uniform_vector(r["v0_seed"], r["n"]), dtype=np.float64)
        keep.append((rowptr, colidx, val, v0))
        reqs[i] = Request(n=len(rowptr) - 1, rowptr=rowptr.ctypes.data, colidx=colidx.ctypes.data,
                          val=None if f32 else val.ctypes.data, val32=val.ctypes.data if f32 else None,
                          nev=r["nev"], ncv=r["ncv"], max_it=0, tol=tol, which=r.get("which", 1),
                          hermitian=int(r.get("hermitian", True)), v0=v0.ctypes.data, want_vectors=int(want_vectors),
                          mixed_tol=r.get("mixed_tol", 0.0), t_arrival=time_scale * r["t_arrival"],
                          twosided=int(r.get("twosided", False)))
        o = dict(eigr=np.zeros(r["nev"]), eigi=np.zeros(r["nev"]), err=np.zeros(r["nev"]))
        if want_vectors:
            o["X"] = np.zeros((r["nev"], r["n"]))
            xs[i] = o["X"].ctypes.data_as(P)
        er[i], ei[i], ee[i] = (o[k].ctypes.data_as(P) for k in ("eigr", "eigi", "err"))
        outs.append(o)
    wall = C.c_double()
    _eps._chk(_eps.lib().SlepcServerRun(server.h, C.c_int(n), reqs, res, er, ei, ee, xs, C.byref(wall)),
              "SlepcServerRun")
    results = []
    for i in range(n):
        q, o = res[i], outs[i]
        k = q.nret
        out = dict(nconv=q.nconv, nret=k, ngood=q.ngood, its=q.its, reason=q.reason, nmatmult=q.nmatmult,
                   gs_passes=q.gs_passes, t_arrival=q.t_arrival, t_admit=q.t_admit, t_done=q.t_done,
                   latency=q.t_done - q.t_arrival, t_host_ds=q.t_host_ds, error=q.error,
                   eigr=o["eigr"][:k], eigi=o["eigi"][:k], err=o["err"][:k])
        if want_vectors:
            out["X"] = o["X"][:k].T.copy()
        results.append(out)
    return results, wall.value

