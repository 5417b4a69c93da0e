"""Batched ensembles: many independent evolve_isothermal runs advanced together.

Each instance is an independent periodic grid with its own shape, operators, integrator,
viscosity/drag, forcing table and end time (the arguments of jathenax/run.py:24-27).
The engine (csrc/jx_batch.hpp) advances all of them with persistent wave-worker kernels,
one RK stage of every active instance per round, with rounds replayed from a hipGraph.
States stay resident in one HBM pool between rounds.  An instance's result depends only
on its own inputs and tile height, never on the rest of the batch.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .device import as_device
from .hydro import run_params


class Ensemble:
    def __init__(self, dtype=torch.float32, max_instances=512, pool_bytes=None, device=0,
                 graph=True, graph_rounds=6):
        if not torch.cuda.is_available():
            raise RuntimeError("Ensemble needs a ROCm GPU (gfx950); none is visible")
        self.dtype = dtype
        self.device = torch.device("cuda", device)
        did = 1 if dtype == torch.float32 else 0
        if pool_bytes is None:
            free, _ = torch.cuda.mem_get_info(device)
            pool_bytes = int(free * 0.5)
        self.pool_bytes = int(pool_bytes)
        h = ctypes.c_void_p()
        lib = _lib.load()
        rc = lib.jx_batch_create(int(device), did, int(max_instances), self.pool_bytes,
                                 ctypes.byref(h))
        if rc != 0:
            raise _lib.JxError("jx_batch_create failed (see stderr)")
        self.handle = h
        self.max_instances = int(max_instances)
        self._keep = {}
        self._shapes = {}
        self.set_graph(graph, graph_rounds)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.jx_batch_destroy(h)
            self.handle = None

    def _check(self, rc, what):
        if rc == 0:
            return
        raw = _lib.load().jx_batch_last_error(self.handle)
        msg = (raw.decode() if raw else "") or what
        if rc == _lib.JX_EKEY:
            raise KeyError(msg)
        if rc == _lib.JX_EINVAL:
            raise ValueError(msg)
        raise _lib.JxError(msg)

    def set_graph(self, on=True, rounds=6):
        self._check(_lib.load().jx_batch_set_graph(self.handle, int(bool(on)), int(rounds)),
                    "set_graph")

    def set_fill(self, fill=2):
        """Tiles per resident wave that each round's tile heights aim for; 0 keeps every
        instance at its own cap.  Results do not depend on it."""
        self._check(_lib.load().jx_batch_set_fill(self.handle, int(fill)), "set_fill")

    def params(self, U0, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, **kw):
        U0 = as_device(U0, dtype=self.dtype, device=self.device)
        return U0, run_params(U0, dx, dy, cs, dfloor, integrator, cfl=cfl, tlim=tlim,
                              nlim=min(int(nlim), 2 ** 31 - 1), **kw)

    def footprint(self, nx, ny, nvar, forced=False):
        p = _lib.RunParams(ny=int(ny), nx=int(nx), nvar=int(nvar))
        return int(_lib.load().jx_batch_footprint(self.handle, ctypes.byref(p), int(forced)))

    def pool_free(self):
        return int(_lib.load().jx_batch_pool_free(self.handle))

    def add(self, U0, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim, *, viscosity=0.0,
            smagorinsky_cs=0.0, drag_rate=0.0, forcing=None, dt_update=1.0, tdriv_start=0.0,
            dedt=0.0, reconstruct="ppm4", rsolver="roe", rows=0, state=None):
        """Register one run (evolve_isothermal's arguments); returns its instance id.
        `state` (a run.LoopState) resumes a checkpointed run: U0 and state as returned by
        checkpoint() continue bit for bit."""
        for name, value in (("viscosity", viscosity), ("smagorinsky_cs", smagorinsky_cs),
                            ("drag_rate", drag_rate), ("dedt", dedt)):
            if not math.isfinite(value) or value < 0.0:
                raise ValueError(f"{name} must be finite and nonnegative")
        U0, p = self.params(U0, dx, dy, cs, dfloor, integrator, cfl, tlim, nlim,
                            reconstruct=reconstruct, rsolver=rsolver, viscosity=viscosity,
                            smagorinsky_cs=smagorinsky_cs, drag_rate=drag_rate,
                            dt_update=dt_update, tdriv_start=tdriv_start, dedt=dedt)
        keep = [U0]
        fptr = None
        if forcing is not None:
            coefficients, basis = forcing
            coef = as_device(coefficients, dtype=self.dtype, device=self.device)
            tabs = [as_device(b, dtype=self.dtype, device=self.device) for b in basis]
            keep += [coef, *tabs]
            f = _lib.Forcing(coefficients=coef.data_ptr(), cos_x=tabs[0].data_ptr(),
                             sin_x=tabs[1].data_ptr(), cos_y=tabs[2].data_ptr(),
                             sin_y=tabs[3].data_ptr(), n_updates=coef.shape[0],
                             n_modes=coef.shape[1])
            fptr = ctypes.byref(f)
        if not rows:
            from . import tuning
            rows = tuning.batch_rows(p.nx, p.ny)
        # The batch runs on its own stream: make the caller's writes visible first.
        torch.cuda.current_stream(self.device).synchronize()
        iid = ctypes.c_int()
        if state is None:
            self._check(_lib.load().jx_batch_add(self.handle, ctypes.byref(p),
                                                 ctypes.c_void_p(U0.data_ptr()), fptr,
                                                 int(rows), ctypes.byref(iid)), "add")
        else:
            s0 = _lib.LoopState(time=float(state.time), prev_dt=float(state.dt),
                                cycle=int(state.cycle), update=int(state.update))
            self._check(_lib.load().jx_batch_add_resume(self.handle, ctypes.byref(p),
                                                        ctypes.c_void_p(U0.data_ptr()), fptr,
                                                        int(rows), ctypes.byref(s0),
                                                        ctypes.byref(iid)), "add")
        self._keep[iid.value] = keep[1:]   # forcing tables must outlive the instance
        self._shapes[iid.value] = tuple(U0.shape)
        return iid.value

    def checkpoint(self, iid):
        """(U copy, run.LoopState) of an instance between rounds, for pre-emption:
        remove() it, and later add(U, ..., state=...) continues it bit for bit."""
        from .run import LoopState
        U = torch.empty(self._shapes[iid], dtype=self.dtype, device=self.device)
        st = _lib.LoopState()
        self._check(_lib.load().jx_batch_checkpoint(self.handle, int(iid),
                                                    ctypes.c_void_p(U.data_ptr()),
                                                    ctypes.byref(st)), "checkpoint")
        return U, LoopState(st.time, st.prev_dt, st.cycle, st.update)

    def remove(self, iid):
        self._check(_lib.load().jx_batch_remove(self.handle, int(iid)), "remove")
        self._keep.pop(iid, None)
        self._shapes.pop(iid, None)

    def run(self, rounds):
        """Run at least `rounds` rounds (one RK stage each); returns the active count."""
        act = ctypes.c_int()
        self._check(_lib.load().jx_batch_run(self.handle, int(rounds), ctypes.byref(act)), "run")
        return act.value

    def launch(self, rounds):
        """Enqueue `rounds` rounds on the engine's stream and return at once; wait() ends
        them.  Work on other streams overlaps them on the GPU."""
        self._check(_lib.load().jx_batch_launch(self.handle, int(rounds)), "launch")

    def wait(self):
        """Wait for the launched rounds; returns the active count."""
        act = ctypes.c_int()
        self._check(_lib.load().jx_batch_wait(self.handle, ctypes.byref(act)), "wait")
        return act.value

    def run_until_done(self, chunk=48, max_rounds=10 ** 9):
        done = 0
        while done < max_rounds:
            act = self.run(chunk)
            done += chunk
            if act == 0:
                return done
        return done

    def poll_launched(self):
        """(done, time, cycle, active) as of the end of the oldest launch() not yet polled:
        waits for that launch only, so a caller that keeps the next launch queued leaves the
        GPU no idle time while it reads the flags."""
        n = self.max_instances
        d, t, c = (ctypes.c_int * n)(), (ctypes.c_double * n)(), (ctypes.c_int * n)()
        act = ctypes.c_int()
        self._check(_lib.load().jx_batch_poll_launched(self.handle, d, t, c, n,
                                                       ctypes.byref(act)), "poll_launched")
        return np.asarray(d[:]), np.asarray(t[:]), np.asarray(c[:]), act.value

    def poll(self):
        n = self.max_instances
        d, t, c = (ctypes.c_int * n)(), (ctypes.c_double * n)(), (ctypes.c_int * n)()
        self._check(_lib.load().jx_batch_poll(self.handle, d, t, c, n), "poll")
        return np.asarray(d[:]), np.asarray(t[:]), np.asarray(c[:])

    def result(self, iid):
        """(U copy, time, last dt, cycle) of an instance."""
        U = torch.empty(self._shapes[iid], dtype=self.dtype, device=self.device)
        t, dt, cyc = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(_lib.load().jx_batch_result(self.handle, int(iid),
                                                ctypes.c_void_p(U.data_ptr()),
                                                ctypes.byref(t), ctypes.byref(dt),
                                                ctypes.byref(cyc)), "result")
        return U, t.value, dt.value, cyc.value
