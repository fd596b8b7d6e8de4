"""Assignment server: exact nearest-centroid queries with continuous batching.

Each query is one distribution (support points and weights); its answer is
what the reference's ``--eval`` mode writes for one object
(``d2_assignment`` -> ``d2_labeling``, clustering.c:324-376, 528-546): the
label of the nearest centroid under the exact transport distance (first
minimum on ties) and that distance.  The work runs in libd2mi.so
(csrc/serve.hip); this module is the Python host side.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Optional, Sequence

import numpy as np

from ._lib import ServeOpts, ServeStats, check, lib, ptr


class Server:
    """Wraps ``d2mi_serve_*`` on the centroids resident in an Engine phase.

    ``plain=True`` serves one query per launch, in arrival order, with one
    group in flight (the one-request-per-launch form)."""

    def __init__(self, engine, phase: int = 0, max_batch: int = 48, streams: int = 6,
                 hbm_bytes: int = 1 << 30, max_m: int = 2048, use_triangle: bool = True,
                 plain: bool = False):
        if plain:
            max_batch, streams = 1, 1
        o = ServeOpts(max_batch, streams, hbm_bytes, max_m, 1 if use_triangle else 0)
        h = C.c_void_p()
        check(lib().d2mi_serve_open(engine._h, phase, C.byref(o), C.byref(h)))
        self._h = h
        self.engine = engine
        self.dim = engine.data.phases[phase].dim

    def submit(self, w: np.ndarray, supp: np.ndarray) -> int:
        w = np.ascontiguousarray(w, np.float64)
        supp = np.ascontiguousarray(supp, np.float64).reshape(len(w), self.dim)
        t = C.c_longlong(0)
        check(lib().d2mi_serve_submit(self._h, len(w), ptr(w), ptr(supp), C.byref(t)))
        return t.value

    def progress(self) -> int:
        p = C.c_int(0)
        check(lib().d2mi_serve_progress(self._h, C.byref(p)))
        return p.value

    def result(self, ticket: int):
        """(label, dist, latency_ms) once done, else None."""
        lab, dist, lat = C.c_int(0), C.c_double(0), C.c_double(0)
        rc = lib().d2mi_serve_result(self._h, ticket, C.byref(lab), C.byref(dist), C.byref(lat))
        if rc < 0:
            check(rc)
        return (lab.value, dist.value, lat.value) if rc == 1 else None

    def drain(self):
        check(lib().d2mi_serve_drain(self._h))

    def stats(self) -> dict:
        st = ServeStats()
        check(lib().d2mi_serve_stats_get(self._h, C.byref(st)))
        return {n: getattr(st, n) for n, _ in ServeStats._fields_}

    def close(self):
        if self._h:
            lib().d2mi_serve_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- drivers
    def assign(self, queries: Sequence[tuple]):
        """Submits every (w, supp) at once and waits: labels, distances."""
        tickets = [self.submit(w, s) for w, s in queries]
        self.drain()
        out = [self.result(t) for t in tickets]
        return (np.array([o[0] for o in out], np.int32), np.array([o[1] for o in out]))

    def replay(self, queries: Sequence[tuple], arrivals_s: np.ndarray, idle_sleep_s: float = 0.0):
        """Open-loop replay: query i is submitted at arrivals_s[i] seconds
        after the start; progress is polled in between.  Returns labels,
        distances, latencies (ms, submission to completion) and the elapsed
        wall time in seconds."""
        n = len(queries)
        order = np.argsort(arrivals_s, kind="stable")
        tickets = np.zeros(n, np.int64)
        labels = np.full(n, -1, np.int32)
        dist = np.zeros(n)
        lat = np.zeros(n)
        done = np.zeros(n, bool)
        t0 = time.perf_counter()
        nxt = 0
        open_ = []
        while nxt < n or open_:
            now = time.perf_counter() - t0
            while nxt < n and arrivals_s[order[nxt]] <= now:
                i = order[nxt]
                tickets[i] = self.submit(*queries[i])
                open_.append(i)
                nxt += 1
            self.progress()
            still = []
            for i in open_:
                r = self.result(int(tickets[i]))
                if r is None:
                    still.append(i)
                else:
                    labels[i], dist[i], lat[i] = r
                    done[i] = True
            open_ = still
            if idle_sleep_s and nxt < n and not open_:
                time.sleep(min(idle_sleep_s, max(0.0, arrivals_s[order[nxt]] - now)))
        return labels, dist, lat, time.perf_counter() - t0
