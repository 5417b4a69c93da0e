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

Failure handling: a request whose time or state is non-finite, or that exceeds its round
budget, is evicted with an error status instead of blocking the batch.

An instance's arithmetic never depends on the other instances (per-instance reductions,
fixed tiles), so a result served in a mix equals the same request served alone.
"""

import math
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
    info: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


class Server:
    def __init__(self, pool_bytes=None, max_instances=256, dtype=torch.float32,
                 rounds_per_poll=6, keep_results=True, graph=True, preempt=True):
        self.ens = Ensemble(dtype=dtype, max_instances=max_instances, pool_bytes=pool_bytes,
                            graph=graph, graph_rounds=rounds_per_poll)
        self.dtype = dtype
        self.rounds_per_poll = rounds_per_poll
        self.keep_results = keep_results
        self.preempt = preempt
        self.queue = deque()
        self.running = {}
        self.finished = []
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

    def _next(self):
        """The highest-priority request that has arrived (the queue is sorted)."""
        now = self.now()
        return next((r for r in self.queue if r.arrival <= now), None)

    def _admit(self):
        while True:
            req = self._next()
            if req is None:
                return
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
            self.running[req.iid] = req

    def _retire(self):
        done, times, cycles = self.ens.poll()
        for iid in list(self.running):
            req = self.running[iid]
            req.rounds += self.rounds_per_poll
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

    def serve(self, requests, timeout=None):
        """Run until every request has finished (or `timeout` seconds elapsed)."""
        for r in requests:
            self.submit(r)
        self.clock0 = time.perf_counter()
        while self.queue or self.running:
            self._admit()
            if self.running:
                self.ens.run(self.rounds_per_poll)
                self._retire()
            elif self.queue:
                wait = min(r.arrival for r in self.queue) - self.now()
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            if timeout is not None and self.now() > timeout:
                break
        return self.finished

    @staticmethod
    def report(finished, wall, zone_stages):
        lat = np.asarray([r.latency for r in finished if r.status == "done"])
        return dict(completed=int(lat.size), errors=sum(r.status == "error" for r in finished),
                    p50_s=float(np.percentile(lat, 50)) if lat.size else None,
                    p99_s=float(np.percentile(lat, 99)) if lat.size else None,
                    wall_s=wall, zone_updates_per_s=zone_stages / wall if wall > 0 else None)
