"""Continuous-batching solve server on one GPU.

Requests (independent CW-FDTD solves of any grid size) arrive over time.  A
request is admitted when its estimated device state fits in the free part of
an HBM budget; its problem is then built on a host worker thread (material
sampling, upload) while the GPU keeps stepping the running batch, and it
joins that batch at the next graph-replay boundary.  Finished solves leave the
batch at once, release their budget and hand their result to a callback.  The
server records per-request latency (arrival to result) and throughput.

Admission is first-come first-served: the queue head waits until it fits
(no overtaking, so no request starves).  At most `max_running` solves are
members of the batch; up to `prep_ahead` more are built ahead on the
workers, so a finished solve is replaced at the next replay boundary.  The
batch steps its members in cache-sized groups in admission order, the
oldest group first, so the device works through requests in arrival order
while the host builds the next ones.  Results do not depend on admission
order or on the other requests in the batch: each member's arithmetic is its
own.
"""
from __future__ import annotations

import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait

import numpy as np

from .batch import Batch

__all__ = ["Request", "Server", "state_bytes"]

BATCH_MAX = 128


def state_bytes(Nx, Ny, Nz, w, slots=1, wavelength=None, d=None):
    """Device bytes of one solver: E and H (48 B/cell), update coefficients
    (24 B) and their compressed index (up to 4 B), the eps copy kept for box
    updates (48 B), result slots (96 B each), the CPML psi slabs (E and H
    sets, 2 components each), the material sampler's staging during the
    build (16 B/cell), and, given the wavelength and cell size, the batch's
    per-step factor table (64 B per step up to the step limit
    500 wavelength/d + the two capture periods) and convergence history."""
    n = Nx * Ny * Nz
    slab = 2 * (Ny * Nz * w[0] + Ny * Nz * w[1] + Nx * Nz * w[2] + Nx * Nz * w[3]
                + Nx * Ny * w[4] + Nx * Ny * w[5])
    total = n * (48 + 24 + 4 + 48 + 96 * slots) + 2 * slab * 8
    if wavelength is not None and d is not None:
        nlambda = wavelength / d
        steps = 500 * nlambda + 4 * np.sqrt(3) * nlambda + 1
        total += int(steps) * 64 + int(steps / max(1.0, 0.75 * np.sqrt(3) * nlambda) + 4) * 24
    return int(total) + (n * 16 if wavelength is not None else 0)


class Request:
    """One solve.  `build` returns a configured FDTD object (materials and
    sources set); `nbytes` is its device-state estimate (state_bytes)."""

    def __init__(self, rid, build, nbytes, adjoint=False, cells=0, meta=None):
        self.rid = rid
        self.build = build
        self.nbytes = int(nbytes)
        self.adjoint = adjoint
        self.cells = cells
        self.meta = meta or {}
        self.t_arrival = None
        self.t_admit = None
        self.t_start = None
        self.t_done = None
        self.info = None
        self.error = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class Server:
    def __init__(self, device=0, hbm_budget=200e9, prep_workers=4, steps_per_graph=32, zc=2,
                 blocks_per_cu=16, graphs_per_poll=4, on_done=None, batch=None, max_running=8,
                 prep_ahead=None, free_hbm=None, hbm_reserve=2 << 30):
        self.batch = batch if batch is not None else Batch(device, steps_per_graph, zc,
                                                           blocks_per_cu)
        self.budget = float(hbm_budget)
        self.used = 0.0
        self.queue = deque()
        self.preparing = []            # (request, future)
        self.running = {}              # member id -> (request, sim)
        self.done = []
        self.pool = ThreadPoolExecutor(max_workers=prep_workers)
        self.graphs_per_poll = graphs_per_poll
        self.on_done = on_done
        self.max_running = min(int(max_running), BATCH_MAX)
        self.prep_ahead = prep_workers if prep_ahead is None else int(prep_ahead)
        # free device memory: the batch's GPU by default (a stand-in batch
        # without the query has none to report)
        if free_hbm is None:
            free_hbm = getattr(self.batch, "free_bytes", None)
        self.free_hbm = free_hbm
        self.hbm_reserve = float(hbm_reserve)
        self.waits_on_free = 0         # admissions deferred by free memory

    # ------------------------------------------------------------ requests
    def submit(self, req, now=None):
        req.t_arrival = time.perf_counter() if now is None else now
        if req.nbytes > self.budget:
            req.error = "request needs %.3g B, more than the whole budget" % req.nbytes
            req.t_done = req.t_arrival
            self.done.append(req)
            return
        self.queue.append(req)

    def _fits_free(self, head):
        """The device has room for `head` now: free memory, less the reserve
        and less what admitted requests still building will allocate."""
        if self.free_hbm is None:
            return True
        pending = sum(r.nbytes for r, f in self.preparing if not f.done())
        if head.nbytes + pending <= self.free_hbm() - self.hbm_reserve:
            return True
        if not self.running and not self.preparing:
            return True            # nothing will free memory: try, and fail alone if it is short
        self.waits_on_free += 1
        return False

    def _admit(self):
        while self.queue:
            head = self.queue[0]
            busy = len(self.running) + len(self.preparing)
            if busy >= self.max_running + self.prep_ahead or self.used + head.nbytes > self.budget:
                break
            if not self._fits_free(head):
                break
            self.queue.popleft()
            self.used += head.nbytes
            head.t_admit = time.perf_counter()
            self.preparing.append((head, self.pool.submit(head.build)))

    def _start_prepared(self):
        keep = []
        for req, fut in self.preparing:      # in admission order
            if not fut.done() or len(self.running) >= self.max_running:
                keep.append((req, fut))
                continue
            try:
                sim = fut.result()
                mid = self.batch.add(sim, adjoint=req.adjoint, postprocess=False)
            except Exception as exc:          # a request that cannot run fails alone
                req.error = repr(exc)
                req.t_done = time.perf_counter()
                self.used -= req.nbytes
                self.done.append(req)
                continue
            req.t_start = time.perf_counter()
            self.running[mid] = (req, sim)
        self.preparing = keep

    def _collect(self, finished_ids):
        now = time.perf_counter()
        for mid in finished_ids:
            req, sim = self.running.pop(mid)
            req.t_done = now
            req.info = self.batch.finished.get(mid)
            self.used -= req.nbytes
            if self.on_done is not None:
                self.on_done(req, sim)
            self.done.append(req)
            del sim

    def poll(self):
        """One scheduling round: admit, start prepared requests, step the
        batch (a few graph replays, returning early when a solve finishes),
        retire finished solves.  Returns the number retired."""
        self._admit()
        self._start_prepared()
        if not self.running:
            if self.preparing:                     # wait for a build to finish
                wait([f for _, f in self.preparing], return_when=FIRST_COMPLETED)
            return 0
        ids = self.batch.run(until="any", max_graphs=self.graphs_per_poll)
        self._collect(ids)
        return len(ids)

    @property
    def idle(self):
        return not (self.queue or self.preparing or self.running)

    def drain(self):
        while not self.idle:
            self.poll()

    def run_trace(self, trace, drain=True, t0=None):
        """Serve `trace` = [(arrival offset in s, Request)] in real time."""
        trace = sorted(trace, key=lambda x: x[0])
        t0 = time.perf_counter() if t0 is None else t0
        k = 0
        while k < len(trace) or (drain and not self.idle):
            now = time.perf_counter()
            while k < len(trace) and t0 + trace[k][0] <= now:
                self.submit(trace[k][1], now=t0 + trace[k][0])
                k += 1
            if self.idle and k < len(trace):
                time.sleep(min(0.01, max(0.0, t0 + trace[k][0] - time.perf_counter())))
                continue
            self.poll()
        return self.stats(t0)

    def stats(self, t0=None):
        ok = [r for r in self.done if r.error is None]
        lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
        out = dict(completed=len(ok), failed=len(self.done) - len(ok))
        if len(lat):
            t_end = max(r.t_done for r in ok)
            t_beg = min(r.t_arrival for r in ok) if t0 is None else t0
            span = max(t_end - t_beg, 1e-9)
            steps = sum(r.info["steps"] * r.cells for r in ok if r.info)
            out.update(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                       mean_s=float(lat.mean()), max_s=float(lat.max()), span_s=span,
                       solves_per_hr=len(ok) * 3600.0 / span,
                       cell_updates_per_s=steps / span)
        return out

    def close(self):
        self.pool.shutdown(wait=True)
        self.batch.close()
