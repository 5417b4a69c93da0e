"""Continuous-batching simulation server.

Requests (evolve_isothermal runs) arrive over time.  The server keeps every admitted
request's state resident in one Ensemble (a single HBM pool) and advances all of them
together, one RK stage per round, so requests that arrive while others are running join
at the next round boundary and their work overlaps on the device.  Admission is by HBM
footprint: a request is admitted when its registers fit the pool's free space and an
instance slot is free; otherwise it waits.  Arrived requests are admitted in priority order
(highest first), then in arrival order.  Finished requests are retired between rounds,
which frees their memory for the queue.

Pre-emption: when the next request cannot be admitted and lower-priority requests are
running, the lowest-priority running request is checkpointed (state and loop state),
removed and requeued; resumed later, it continues bit for bit, so pre-emption changes only
latency.

Large requests: a request of at least `big_zones` cells runs on the single-grid fused
kernel (the device-resident loop, in chunks of `big_chunk` steps) on a stream of its own,
while the batch's rounds run on the engine's stream: one grid of that size fills the GPU
better alone than as tiles among others. By default each large request runs in a worker
thread (the library's calls release the interpreter lock), so the server loop keeps
launching and retiring batch rounds while the large request's chunks run, and two large
requests run at once (a grid of 1024^2 to 1448^2 cells leaves much of the GPU idle); with
`big_threads=False` the server loop runs one chunk of every large request per iteration
itself, one large request at a time. At most `max_big` run at once; the rest wait without
blocking smaller requests. A request's path depends only on its size, so its result
is the same alone as in any mix. In FP32 the two paths agree to round-off (the kernels are
separate code that may fuse different products into adds, and forced runs reduce their kick
sums in different orders).

Failure handling: a request whose time or state is non-finite, or that exceeds its round
budget, is evicted with an error status instead of blocking the batch.

An instance's arithmetic never depends on the other instances (per-instance reductions,
fixed tiles), so a result served in a mix equals the same request served alone.
"""

import math
import queue as _queue
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from .ensemble import Ensemble


@dataclass
class Request:
    rid: int
    spec: dict
    arrival: float                 # seconds after the server clock starts
    rows: int = 0
    max_rounds: int = 10 ** 9
    admitted: float = math.nan
    finished: float = math.nan
    status: str = "queued"         # queued, running, done, error
    result: object = None          # (U on device, time, dt, cycle)
    iid: int = -1
    rounds: int = 0
    priority: int = 0              # larger is more urgent
    checkpoint: object = None      # (U on device, run.LoopState) while pre-empted
    preemptions: int = 0
    first_launch: int = 0          # the first batch launch it takes part in (batch path)
    info: dict = field(default_factory=dict)
    path: str = "batch"            # "batch" or "grid" (single-grid kernel, large requests)
    state: object = None           # grid path: (U on device, run.LoopState) between chunks

    @property
    def latency(self):
        return self.finished - self.arrival


class Server:
    def __init__(self, pool_bytes=None, max_instances=256, dtype=torch.float32,
                 rounds_per_poll=6, keep_results=True, graph=True, preempt=True,
                 big_zones=1 << 20, big_chunk=None, max_big=None, big_rounds=4,
                 big_threads=True):
        self.ens = Ensemble(dtype=dtype, max_instances=max_instances, pool_bytes=pool_bytes,
                            graph=graph, graph_rounds=rounds_per_poll)
        self.dtype = dtype
        self.rounds_per_poll = rounds_per_poll
        self.keep_results = keep_results
        self.preempt = preempt
        self.big_zones = big_zones          # 0: every request takes the batch path
        self.big_threads = big_threads
        # Steps per chunk of a large request: a worker thread checks the request (finished,
        # non-finite, over budget, server stopping) between chunks; the server loop's own
        # chunks are shorter, since the loop waits for them.
        self.big_chunk = big_chunk or (256 if big_threads else 16)
        self.max_big = max_big or (2 if big_threads else 1)
        # Without worker threads, each server iteration launches big_rounds times as many
        # batch rounds while large requests run, so the batch keeps the GPU busy during the
        # large requests' chunks.
        self.big_rounds = big_rounds
        self.queue = deque()
        self.running = {}
        self.grid_running = {}
        self.grid_stream = torch.cuda.Stream()
        # Worker threads: one stream per slot (each stream has its own library context and
        # scratch, reused by the slot's later requests); finished requests come back through
        # grid_done.
        self.grid_streams = ([torch.cuda.Stream() for _ in range(self.max_big)]
                             if big_threads else [])
        self.grid_done = _queue.SimpleQueue()
        self.workers = {}
        self.stop = threading.Event()
        self.finished = []
        # Batch launches issued so far; an admitted request takes part in launches from
        # index first_launch on, so a loop-state snapshot of an earlier launch (which may show
        # the slot's previous occupant) is not applied to it.
        self.launches = 0
        self.clock0 = None

    def now(self):
        return time.perf_counter() - self.clock0

    def submit(self, req):
        """Queue a request: by priority (highest first), then arrival."""
        items = list(self.queue) + [req]
        items.sort(key=lambda r: (-r.priority, r.arrival))
        self.queue = deque(items)

    def _fits(self, req):
        return (len(self.running) < self.ens.max_instances
                and self._footprint(req.spec) <= self.ens.pool_free())

    def _preempt_for(self, req):
        """Checkpoint and requeue lower-priority running requests until `req` fits."""
        victims = sorted((r for r in self.running.values() if r.priority < req.priority),
                         key=lambda r: (r.priority, -r.admitted))
        for v in victims:
            if self._fits(req):
                break
            v.checkpoint = self.ens.checkpoint(v.iid)
            self.ens.remove(v.iid)
            del self.running[v.iid]
            v.iid, v.status = -1, "queued"
            v.preemptions += 1
            self.submit(v)
        return self._fits(req)

    def _footprint(self, spec):
        ny, nx, nv = spec["U"].shape
        return self.ens.footprint(nx, ny, nv, forced="forcing" in spec["kw"]
                                  and spec["kw"]["forcing"] is not None)

    def _is_big(self, req):
        ny, nx = req.spec["U"].shape[:2]
        return bool(self.big_zones) and nx * ny >= self.big_zones

    def _next(self):
        """The highest-priority arrived request (the queue is sorted) that has a path free:
        large requests wait for a single-grid slot without blocking the others."""
        now = self.now()
        grid_full = len(self.grid_running) >= self.max_big
        return next((r for r in self.queue if r.arrival <= now
                     and not (grid_full and self._is_big(r))), None)

    def _admit_grid(self, req):
        self.queue.remove(req)
        s = req.spec
        if req.checkpoint is not None:
            req.state, req.checkpoint = req.checkpoint, None
        else:
            req.state = (torch.as_tensor(s["U"], dtype=self.dtype, device="cuda"), None)
        req.path = "grid"
        if math.isnan(req.admitted):
            req.admitted = self.now()
        req.status = "running"
        self.grid_running[req.rid] = req
        if self.big_threads:
            self._start_worker(req)

    def _start_worker(self, req):
        # The state was made on this thread's stream; the worker's stream waits for it.
        ready = torch.cuda.Event()
        ready.record()
        busy = {w[1] for w in self.workers.values()}
        slot = next(k for k in range(self.max_big) if k not in busy)
        th = threading.Thread(target=self._grid_worker, args=(req, slot, ready),
                              name=f"grid-{req.rid}", daemon=True)
        self.workers[req.rid] = (th, slot)
        th.start()

    def _admit(self):
        while True:
            req = self._next()
            if req is None:
                return
            if self._is_big(req):
                self._admit_grid(req)
                continue
            if not self._fits(req):
                if self.preempt and self._preempt_for(req):
                    pass
                elif not self.running:
                    # Cannot ever fit: reject rather than stall the queue.
                    self.queue.remove(req)
                    req.status, req.finished = "error", self.now()
                    req.info["error"] = "request exceeds the HBM pool"
                    self.finished.append(req)
                    continue
                else:
                    return
            self.queue.remove(req)
            s = req.spec
            if req.checkpoint is not None:
                U0, state = req.checkpoint
                req.checkpoint = None
            else:
                U0, state = torch.as_tensor(s["U"], dtype=self.dtype, device="cuda"), None
            req.iid = self.ens.add(U0, s["dx"], s["dy"], s["cs"], s["dfloor"], s["integrator"],
                                   s["cfl"], s["tlim"], s["nlim"], rows=req.rows, state=state,
                                   **s["kw"])
            if math.isnan(req.admitted):
                req.admitted = self.now()
            req.status = "running"
            req.first_launch = self.launches
            self.running[req.iid] = req

    def _retire(self, rounds, launch, snapshot=None):
        """Retire the requests that the loop state as of the end of batch launch `launch`
        (`snapshot`: done, time, cycle; None: poll the current state) shows finished."""
        done, times, cycles = snapshot if snapshot is not None else self.ens.poll()
        for iid in list(self.running):
            req = self.running[iid]
            if req.first_launch > launch:
                continue   # admitted after that launch: the snapshot predates it
            req.rounds += rounds
            bad = not math.isfinite(times[iid])
            if done[iid] == 1 or bad or req.rounds > req.max_rounds:
                res = self.ens.result(iid)
                req.finished = self.now()
                req.info["cycles"] = int(res[3])
                if bad or not torch.isfinite(res[0]).all():
                    req.status, req.info["error"] = "error", "non-finite state"
                elif done[iid] != 1:
                    req.status, req.info["error"] = "error", "round budget exceeded"
                else:
                    req.status = "done"
                req.result = res if self.keep_results else None
                self.ens.remove(iid)
                del self.running[iid]
                self.finished.append(req)

    def _grid_finish(self, req, U, st):
        """True (and the request closed) when a large request is done, non-finite or over
        its budget."""
        s = req.spec
        bad = not math.isfinite(st.time)
        if not bad and st.time < s["tlim"] and st.cycle < s["nlim"] \
                and req.rounds <= req.max_rounds:
            return False
        req.finished = self.now()
        req.info["cycles"] = int(st.cycle)
        if bad or not torch.isfinite(U).all():
            req.status, req.info["error"] = "error", "non-finite state"
        elif st.time < s["tlim"] and st.cycle < s["nlim"]:
            req.status, req.info["error"] = "error", "round budget exceeded"
        else:
            req.status = "done"
        req.result = (U, st.time, st.dt, st.cycle) if self.keep_results else None
        req.state = None
        return True

    def _grid_worker(self, req, slot, ready):
        """A large request in chunks of big_chunk steps on the slot's stream, until it is
        closed or the server stops (the request then keeps its state and stays running)."""
        from .run import evolve_resumable
        stream = self.grid_streams[slot]
        try:
            with torch.cuda.stream(stream):
                stream.wait_event(ready)
                s = req.spec
                U, st = req.state
                while not self.stop.is_set():
                    target = min(s["nlim"], (st.cycle if st else 0) + self.big_chunk)
                    U, st = evolve_resumable(U, s["dx"], s["dy"], s["cs"], s["dfloor"],
                                             s["integrator"], s["cfl"], s["tlim"], target,
                                             state=st, **s["kw"])
                    req.state = (U, st)
                    req.rounds += self.big_chunk
                    if self._grid_finish(req, U, st):
                        break
                stream.synchronize()
        except Exception as e:   # reported as the request's error, not the server's
            req.finished = self.now()
            req.status, req.info["error"] = "error", f"{type(e).__name__}: {e}"
            req.state = None
        self.grid_done.put(req)

    def _collect_grid(self):
        """Close the large requests whose worker has ended."""
        while True:
            try:
                req = self.grid_done.get_nowait()
            except _queue.Empty:
                return
            th, _ = self.workers.pop(req.rid)
            th.join()
            if req.status != "running":
                del self.grid_running[req.rid]
                self.finished.append(req)

    def _step_grid(self):
        """One chunk of steps for every large request, on the grid stream (the batch's
        launched rounds run meanwhile on the engine's stream)."""
        from .run import evolve_resumable
        self.grid_stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self.grid_stream):
            for rid in list(self.grid_running):
                req = self.grid_running[rid]
                s = req.spec
                U, st = req.state
                target = min(s["nlim"], (st.cycle if st else 0) + self.big_chunk)
                U, st = evolve_resumable(U, s["dx"], s["dy"], s["cs"], s["dfloor"],
                                         s["integrator"], s["cfl"], s["tlim"], target, state=st,
                                         **s["kw"])
                req.state = (U, st)
                req.rounds += self.big_chunk
                if self._grid_finish(req, U, st):
                    del self.grid_running[rid]
                    self.finished.append(req)
        self.grid_stream.synchronize()

    def serve(self, requests, timeout=None):
        """Run until every request has finished (or `timeout` seconds elapsed)."""
        for r in requests:
            self.submit(r)
        self.clock0 = time.perf_counter()
        self.stop.clear()
        threaded = self.big_threads
        if threaded:
            # Large requests a timed-out call left running continue where they stopped.
            for req in self.grid_running.values():
                if req.rid not in self.workers:
                    self._start_worker(req)
        # Batch launches not yet read: (index, rounds).  While requests run, the next launch
        # is queued before the loop state of the previous one is read (poll_launched waits
        # for that launch only), so the batch's rounds run back to back; adding or removing
        # an instance waits for the launches in flight.
        inflight = deque()
        try:
            while self.queue or self.running or self.grid_running or inflight:
                self._admit()
                if threaded:
                    self._collect_grid()
                busy = self.grid_running and not threaded
                rounds = self.rounds_per_poll * (self.big_rounds if busy else 1)
                if self.running:
                    self.ens.launch(rounds)
                    inflight.append((self.launches, rounds))
                    self.launches += 1
                if busy:
                    self._step_grid()
                if len(inflight) > 1:
                    done, times, cycles, _ = self.ens.poll_launched()
                    launch, n = inflight.popleft()
                    self._retire(n, launch, (done, times, cycles))
                elif inflight and not self.running:
                    self.ens.wait()
                    launch, n = inflight.popleft()
                    self._retire(n, launch)
                elif inflight:
                    pass   # the first launch of a run of rounds: read it next iteration
                elif self.grid_running and threaded:
                    time.sleep(0.0005)   # only large requests run: wait for their workers
                elif not self.grid_running and self.queue:
                    wait = min(r.arrival for r in self.queue) - self.now()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                if timeout is not None and self.now() > timeout:
                    break
        finally:
            if inflight:   # a timed-out call: settle the launches in flight
                self.ens.wait()
                while inflight:
                    launch, n = inflight.popleft()
                    self._retire(n, launch)
            # Workers stop between chunks; a request stopped early stays running with its
            # state.
            self.stop.set()
            for th, _ in list(self.workers.values()):
                th.join()
            if threaded:
                self._collect_grid()
        return self.finished

    @staticmethod
    def report(finished, wall, zone_stages):
        """Completion latency (finish - arrival) and its two parts, the time queued before
        admission (admitted - arrival) and the time in service (finish - admitted), as
        p50 / p99 over the completed requests."""
        done = [r for r in finished if r.status == "done"]
        lat = np.asarray([r.latency for r in done])
        queue = np.asarray([r.admitted - r.arrival for r in done])
        service = np.asarray([r.finished - r.admitted for r in done])

        def pct(x, q):
            return float(np.percentile(x, q)) if x.size else None

        return dict(completed=int(lat.size), errors=sum(r.status == "error" for r in finished),
                    p50_s=pct(lat, 50), p99_s=pct(lat, 99),
                    queue_p50_s=pct(queue, 50), queue_p99_s=pct(queue, 99),
                    service_p50_s=pct(service, 50), service_p99_s=pct(service, 99),
                    wall_s=wall, zone_updates_per_s=zone_stages / wall if wall > 0 else None)
