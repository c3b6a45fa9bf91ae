"""Serving channel simulations on one GPU with continuous batching.

Requests (a channel configuration, a fixed time step, a number of SMR91
steps and an initial state) arrive open loop; the server admits each one as
soon as the HBM budget allows, keeps its state and factorizations resident
on the device, advances every live request by one step per iteration in one
ensemble step (members' explicit stages on a few streams, every member's
implicit solves in the grouped mixed precision solver), and retires a
request when its steps are done.  A request admitted while others run
joins their batch at the next iteration.

The reference's driver advances one case per process and step loop
(apps/perfect/main_advance.cpp, suzerain/support/driver_base.cpp:488+);
here many cases share the device.
"""
import math
import time

import numpy as np

from .ensemble import Ensemble, member_bytes, default_dt

__all__ = ["Request", "Server", "poisson_trace", "Fp64Server"]


class Request:
    def __init__(self, rid, cfg, steps, arrival, init, dt=None):
        self.rid = rid
        self.cfg = cfg
        self.steps = steps
        self.arrival = arrival          # seconds after the trace start
        self.init = init                # callable -> device state tensor
        self.dt = default_dt(cfg) if dt is None else dt
        self.admitted = None
        self.finished = None
        self.done = 0
        self.result = None              # host copy of the final state
        self.mid = None


def poisson_trace(rate, count, make, seed=1):
    """count requests with exponential inter-arrival times of mean 1/rate;
    make(i, rng) -> (cfg, steps, init)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        cfg, steps, init = make(i, rng)
        out.append(Request(i, cfg, steps, t, init))
    return out


class Server:
    """Continuous batching over an Ensemble within an HBM budget."""

    def __init__(self, budget_bytes=240e9, solver="zcgbsvx,reuse=true,siter=3,"
                 "diter=5,tolsc=0.1", graph=True, nstreams=1, max_streams=2,
                 keep_results=True, max_live=None):
        import torch
        self.budget = budget_bytes
        self.max_live = max_live
        self.ens = Ensemble([], solver=solver, graph=graph, nstreams=nstreams,
                            max_streams=max_streams)
        self.keep = keep_results
        self.stream = torch.cuda.current_stream()

    def resident(self):
        return self.ens.bytes()

    def fits(self, req):
        return self.resident() + member_bytes(req.cfg) <= self.budget

    def run(self, requests, realtime=True, log=None):
        """Serve the trace; returns a report dict.  realtime: arrivals follow
        the wall clock (idle gaps are slept); else they are admitted as soon
        as memory allows (a closed loop over the same requests)."""
        import torch
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, live, done = [], [], []
        sub_ms = []                     # per-substep device time samples
        t0 = time.perf_counter()
        substeps = 0
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        while pending or queue or live:
            now = time.perf_counter() - t0
            while pending and (pending[0].arrival <= now or not realtime):
                queue.append(pending.pop(0))
            # admission in arrival order, as memory frees up
            while queue and self.fits(queue[0]) and (
                    self.max_live is None or len(live) < self.max_live):
                r = queue.pop(0)
                r.mid = self.ens.add(r.cfg, r.dt)
                self.ens.set_state(self.ens.position(r.mid), r.init())
                r.admitted = time.perf_counter() - t0
                live.append(r)
            if queue and not live:
                raise MemoryError(f"request {queue[0].rid} does not fit the "
                                  f"HBM budget alone")
            if not live:
                if pending and realtime:
                    time.sleep(max(0.0, pending[0].arrival -
                                   (time.perf_counter() - t0)))
                continue
            ev0.record(self.stream)
            self.ens.step(1)
            ev1.record(self.stream)
            ev1.synchronize()
            ms = ev0.elapsed_time(ev1)
            sub_ms.extend([ms / 3.0] * 3)
            substeps += 3 * len(live)
            now = time.perf_counter() - t0
            still = []
            for r in live:
                r.done += 1
                if r.done >= r.steps:
                    if self.keep:
                        r.result = self.ens.get_state(
                            self.ens.position(r.mid)).cpu().numpy()
                    self.ens.remove(r.mid)
                    r.finished = time.perf_counter() - t0
                    done.append(r)
                else:
                    still.append(r)
            live = still
            if log:
                log(now, len(live), len(queue))
        wall = time.perf_counter() - t0
        self.ens.check()
        lat = [r.finished - r.arrival for r in done]
        return {
            "requests": len(done),
            "wall_s": round(wall, 3),
            "member_substeps": substeps,
            "substeps_per_s": round(substeps / wall, 2),
            "substep_ms": _pct(sub_ms),
            "completion_s": _pct(lat),
            "queue_wait_s": _pct([r.admitted - r.arrival for r in done]),
        }


class Fp64Server:
    """The plain FP64 drop-in under the same trace: every request is its
    own channel (zgbsv, refactored every substep), the live ones stepped
    round robin, admitted within the same budget."""

    def __init__(self, budget_bytes=240e9):
        self.budget = budget_bytes

    @staticmethod
    def bytes_of(cfg):
        Nx, Ny, Nz, k = cfg["Nx"], cfg["Ny"], cfg["Nz"], cfg.get("k", 8)
        dNx, dNz = 3 * Nx // 2, 3 * Nz // 2
        F = dNz * (dNx // 2 + 1) * Ny
        nsys = ((Nx - 1) // 2 + 1) * (2 * ((Nz - 1) // 2) + 1)
        KL = 5 * (k - 1) - 1
        N = 5 * Ny
        lu = nsys * N * (3 * KL + 1) * 16
        return lu + (2 + 5 + 41 + 46) * F * 16 + 46 * Ny * dNz * dNx * 8

    def run(self, requests, realtime=True):
        import torch
        from .channel import Channel
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, live, done = [], [], []
        resident = 0
        sub_ms = []
        substeps = 0
        t0 = time.perf_counter()
        st = torch.cuda.current_stream()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        while pending or queue or live:
            now = time.perf_counter() - t0
            while pending and (pending[0].arrival <= now or not realtime):
                queue.append(pending.pop(0))
            while queue and resident + self.bytes_of(queue[0].cfg) <= self.budget:
                r = queue.pop(0)
                cfg = dict(r.cfg)
                r.ch = Channel(**cfg, solver="zgbsv")
                r.ch.set_state(r.init())
                resident += self.bytes_of(r.cfg)
                r.admitted = time.perf_counter() - t0
                live.append(r)
            if not live:
                if pending and realtime:
                    time.sleep(max(0.0, pending[0].arrival -
                                   (time.perf_counter() - t0)))
                continue
            still = []
            for r in live:
                ev0.record(st)
                r.ch.step(max_dt=r.dt)
                ev1.record(st)
                ev1.synchronize()
                sub_ms.extend([ev0.elapsed_time(ev1) / 3.0] * 3)
                substeps += 3
                r.done += 1
                if r.done >= r.steps:
                    r.result = r.ch.get_state(0).cpu().numpy()
                    del r.ch
                    resident -= self.bytes_of(r.cfg)
                    r.finished = time.perf_counter() - t0
                    done.append(r)
                else:
                    still.append(r)
            live = still
        wall = time.perf_counter() - t0
        lat = [r.finished - r.arrival for r in done]
        return {
            "requests": len(done),
            "wall_s": round(wall, 3),
            "member_substeps": substeps,
            "substeps_per_s": round(substeps / wall, 2),
            "substep_ms": _pct(sub_ms),
            "completion_s": _pct(lat),
        }


def _pct(v):
    if not v:
        return None
    a = np.asarray(v, dtype=float)
    return {"p50": round(float(np.percentile(a, 50)), 4),
            "p99": round(float(np.percentile(a, 99)), 4),
            "mean": round(float(a.mean()), 4), "count": int(a.size)}
