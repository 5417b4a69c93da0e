"""Continuous-batching simulation server.

Requests (evolve_isothermal runs) arrive over time.  The server keeps every admitted
request's state resident in one Ensemble (a single HBM pool) and advances all of them
together, one RK stage per round, so requests that arrive while others are running join
at the next round boundary and their work overlaps on the device.  Admission is by HBM
footprint: a request is admitted when its registers fit the pool's free space and an
instance slot is free; otherwise it waits in FIFO order.  Finished requests are retired
between rounds, which frees their memory for the queue.

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
    info: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


class Server:
    def __init__(self, pool_bytes=None, max_instances=256, dtype=torch.float32,
                 rounds_per_poll=6, keep_results=True, graph=True):
        self.ens = Ensemble(dtype=dtype, max_instances=max_instances, pool_bytes=pool_bytes,
                            graph=graph, graph_rounds=rounds_per_poll)
        self.dtype = dtype
        self.rounds_per_poll = rounds_per_poll
        self.keep_results = keep_results
        self.queue = deque()
        self.running = {}
        self.finished = []
        self.clock0 = None

    def now(self):
        return time.perf_counter() - self.clock0

    def submit(self, req):
        self.queue.append(req)

    def _footprint(self, spec):
        ny, nx, nv = spec["U"].shape
        return self.ens.footprint(nx, ny, nv, forced="forcing" in spec["kw"]
                                  and spec["kw"]["forcing"] is not None)

    def _admit(self):
        while self.queue and self.queue[0].arrival <= self.now():
            req = self.queue[0]
            if len(self.running) >= self.ens.max_instances:
                return
            if self._footprint(req.spec) > self.ens.pool_free():
                if not self.running:
                    # Cannot ever fit: reject rather than stall the queue.
                    self.queue.popleft()
                    req.status, req.finished = "error", self.now()
                    req.info["error"] = "request exceeds the HBM pool"
                    self.finished.append(req)
                    continue
                return
            self.queue.popleft()
            s = req.spec
            req.iid = self.ens.add(torch.as_tensor(s["U"], dtype=self.dtype, device="cuda"),
                                   s["dx"], s["dy"], s["cs"], s["dfloor"], s["integrator"],
                                   s["cfl"], s["tlim"], s["nlim"], rows=req.rows, **s["kw"])
            req.admitted, req.status = self.now(), "running"
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
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        self.clock0 = time.perf_counter()
        while self.queue or self.running:
            self._admit()
            if self.running:
                self.ens.run(self.rounds_per_poll)
                self._retire()
            elif self.queue:
                wait = self.queue[0].arrival - self.now()
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
