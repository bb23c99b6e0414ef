"""Serving form: Python handle on the engine's continuous-batching server, and an open-loop
Poisson load generator that reports atoms/s and p50/p99 request latency.

Latency of a request = completion time - scheduled arrival time, both on the engine's
monotonic clock (nnp_server_now), so queueing before submission counts against the server.
"""
from __future__ import annotations

import ctypes
import threading
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import nnp as _n
from .synth import Structure


class _Cfg(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint32), ("lanes", ctypes.c_int32), ("hbm_budget", ctypes.c_int64),
                ("max_batch_atoms", ctypes.c_int32), ("max_batch_structures", ctypes.c_int32)]


_bound = False


def _lib():
    global _bound
    L = _n.lib()
    if not _bound:
        vp = ctypes.c_void_p
        L.nnp_server_create.argtypes = [vp, ctypes.POINTER(_Cfg), ctypes.POINTER(vp)]
        L.nnp_server_submit.argtypes = [vp, ctypes.POINTER(_n._Struc), ctypes.POINTER(_n._Result)]
        L.nnp_server_submit.restype = ctypes.c_int64
        L.nnp_server_poll.argtypes = [vp, ctypes.c_int64]
        L.nnp_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
        L.nnp_server_now.restype = ctypes.c_double
        L.nnp_server_stats.argtypes = [vp, vp, ctypes.c_int]
        L.nnp_server_destroy.argtypes = [vp]
        _bound = True
    return L


def now() -> float:
    return _lib().nnp_server_now()


class Pending:
    """A submitted request: keeps its input and output buffers alive until waited for."""

    def __init__(self, server: "Server", s: Structure):
        self.pin = _n._Pinned(s)
        self.eatom = np.zeros(s.natoms)
        self.force = np.zeros((s.natoms, 3))
        self.res = _n._Result(0.0, self.eatom.ctypes.data, self.force.ctypes.data)
        self.id = _n._check(_lib().nnp_server_submit(server._h, ctypes.byref(self.pin.c_struct),
                                                    ctypes.byref(self.res)))
        self.server = server
        self.t_submit = 0.0
        self.t_done = 0.0

    def wait(self):
        ts, td = ctypes.c_double(), ctypes.c_double()
        _n._check(_lib().nnp_server_wait(self.server._h, self.id, ctypes.byref(ts), ctypes.byref(td)))
        self.t_submit, self.t_done = ts.value, td.value
        return self.res.energy, self.eatom, self.force


class Server:
    def __init__(self, model: _n.NNP, flags: int, lanes: int = 3, hbm_budget: int = 16 << 30,
                 max_batch_atoms: int = 1 << 18, max_batch_structures: int = 4096):
        cfg = _Cfg(flags, lanes, hbm_budget, max_batch_atoms, max_batch_structures)
        h = ctypes.c_void_p()
        _n._check(_lib().nnp_server_create(model._h, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.model = model

    def submit(self, s: Structure) -> Pending:
        return Pending(self, s)

    def stats(self) -> dict:
        out = np.zeros(6, dtype=np.int64)
        _n._check(_lib().nnp_server_stats(self._h, out.ctypes.data, 6))
        return dict(zip(["completed", "batches", "max_concurrent_batches", "workspace_bytes", "peak_bytes",
                         "queued"], (int(v) for v in out)))

    def close(self):
        if getattr(self, "_h", None):
            _lib().nnp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class StreamResult:
    requests: int
    atoms: int
    wall_s: float
    atoms_per_s: float
    p50_ms: float
    p99_ms: float
    mean_ms: float
    stats: dict
    results: Optional[list] = None


def run_stream(server: Server, structures: Sequence[Structure], arrivals: Sequence[float],
               keep_results: bool = False) -> StreamResult:
    """Open loop: request k is submitted at t0 + arrivals[k] whatever the server's state; a
    collector thread waits for completions in submission order."""
    n = len(structures)
    pend: List[Optional[Pending]] = [None] * n
    done_lat = np.zeros(n)
    results = [None] * n if keep_results else None
    ready = threading.Semaphore(0)
    t0 = now() + 0.05

    def collect():
        for k in range(n):
            ready.acquire()
            E, Ea, F = pend[k].wait()
            done_lat[k] = pend[k].t_done - (t0 + arrivals[k])
            if keep_results:
                results[k] = (E, Ea.copy(), F.copy())
            pend[k] = None  # release buffers

    th = threading.Thread(target=collect)
    th.start()
    for k in range(n):
        delay = t0 + arrivals[k] - now()
        if delay > 0:
            time.sleep(delay)
        pend[k] = server.submit(structures[k])
        ready.release()
    th.join()
    wall = now() - t0
    atoms = int(sum(s.natoms for s in structures))
    lat = done_lat * 1e3
    return StreamResult(n, atoms, wall, atoms / wall, float(np.percentile(lat, 50)), float(np.percentile(lat, 99)),
                        float(lat.mean()), server.stats(), results)
