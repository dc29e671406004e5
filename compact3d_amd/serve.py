"""Serving many channel instances on one GPU: a continuous-batching request
loop over resident states.

A request asks for `steps` time steps of one channel (grid, Reynolds
number, optional Smagorinsky + wall model, precision) from a seeded initial
field.  The loop admits waiting requests while their HBM footprint
(`c3d_footprint_bytes`) fits the budget and a slot is free, keeps every
admitted state resident, and advances the live instances as a rolling
pipeline: every live instance always has a chunk of up to `chunk` steps
queued on its own HIP stream (`c3d_advance`, replaying captured graphs).
The loop waits for the instance whose chunk was queued first, reads its
DIVMAX and CFL, queues its next chunk at once (or retires it), and admits
any request that has arrived and now fits before waiting for the next one.
Work of different instances therefore overlaps on the device, a request
that arrives while others run is admitted as soon as a slot and memory are
free (at the next chunk boundary of any instance, not after a whole round),
and a finished instance frees its memory and slot at once.  Each instance's
arithmetic does not depend on what else is resident, so a result is the
same alone and in the mix.

Two policies: "fifo" admits in arrival order and keeps a chunk queued for
every resident instance (the device is shared by all of them, so every
request's latency grows with the number resident); "srpt" (shortest
remaining work first, work = cells x steps) admits the smallest waiting
request first and keeps chunks queued only for the `max_active` resident
instances with the least remaining work, re-chosen at every chunk boundary
(a newly admitted short request overtakes long ones within one chunk).
SRPT minimises the mean latency of a given set of jobs; the last job ends
when the total work is done under either policy.  `max_active` defaults to
8, twice the 4 hardware queues a process gets by default: more streams than
that do not run side by side, and the spare ones keep queued work ahead of
the loop's chunk boundaries.

Failure detection: an instance whose DIVMAX or CFLM comes back NaN or
infinite (c3d_sync's C3D_ERR_NONFINITE) has diverged; only that request is
marked failed (`Request.error`) and evicted, and the others continue.

The reference runs one channel per process and has no counterpart; this is
the ensemble/serving driver over the library's C-ABI.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

PI = math.pi


def re_from_re_tau(re_tau: float) -> float:
    """RE in the reference's units (centreline velocity 1, half height 1)
    for a friction Reynolds number, via Dean's correlation
    Re_tau = 0.09 Re_b^0.88 (Re_b on the full height and bulk velocity) and
    RE = 0.75 Re_b (compact3d.f:627, :656)."""
    re_b = (re_tau / 0.09) ** (1.0 / 0.88)
    return 0.75 * re_b


@dataclass
class Request:
    rid: int
    grid: tuple          # (N1M, N2, N3M)
    re: float
    steps: int
    seed: int = 0
    precision: int = 0   # 0 FP64, 1 mixed
    les: bool = False    # Smagorinsky SGS (FILES=2) + the log-law wall model at both walls (WMODEL=6)
    dt: float = 5e-4
    cks: float = 1.5
    t_arrive: float = 0.0
    # filled in by the loop
    t_admit: float = float("nan")
    t_done: float = float("nan")
    result: dict = field(default_factory=dict)
    error: str = ""       # set when the instance diverged (NaN/Inf DIVMAX or CFL)

    def config(self):
        n1m, n2, n3m = self.grid
        return dict(n1=n1m + 1, n2=n2, n3=n3m + 1, re=self.re, alx=2 * PI, aly=2.0, alz=PI, dt=self.dt,
                    files=2.0 if self.les else 0.0, wmodel=6.0 if self.les else 0.0, cks=self.cks)

    @property
    def pts(self):
        n1m, n2, n3m = self.grid
        return n1m * (n2 - 1) * n3m


class FieldCache:
    """Seeded initial fields, one per (grid, seed): the divergence-free
    perturbed Poiseuille field of compact3d_amd.initfield on the grid the
    library builds."""

    def __init__(self):
        self._f = {}

    def get(self, req: Request, mesh):
        from .initfield import perturbed_poiseuille
        key = (req.grid, req.seed, req.les, req.cks)
        if key not in self._f:
            cfg = req.config()
            self._f[key] = perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], mesh["Y"], mesh["DX1"],
                                                mesh["DX3"], mesh["ALY"], seed=req.seed)
        return self._f[key]


def _channel_factory(req: Request):
    from . import c3d
    return c3d.Channel(**req.config(), precision=req.precision)


def _footprint(req: Request) -> int:
    from . import c3d
    return c3d.footprint_bytes(**req.config(), precision=req.precision)


class Server:
    """Continuous-batching loop.  `make_channel(req)` and `footprint(req)`
    default to the library; tests may pass stand-ins."""

    def __init__(self, budget_bytes: int, max_live: int = 64, chunk: int = 10, use_graph: bool = True,
                 make_channel=None, footprint=None, fields: FieldCache | None = None, keep_results: bool = False,
                 clock=time.perf_counter, sleep=time.sleep, policy: str = "fifo", max_active: int = 8):
        if policy not in ("fifo", "srpt"):
            raise ValueError(f"unknown policy {policy!r} (fifo or srpt)")
        if max_active < 1:
            raise ValueError("max_active must be at least 1")
        self.policy = policy
        self.max_active = int(max_active)
        self.budget = int(budget_bytes)
        self.max_live = int(max_live)
        self.chunk = int(chunk)
        self.use_graph = use_graph
        self.make_channel = make_channel or _channel_factory
        self.footprint = footprint or _footprint
        self.fields = fields or FieldCache()
        self.keep_results = keep_results
        self.clock = clock
        self.sleep = sleep
        self.peak_bytes = 0

    def _admit(self, req, now):
        ch = self.make_channel(req)
        U, P = self.fields.get(req, ch.mesh())
        ch.upload(U, P)
        req.t_admit = now
        return dict(req=req, ch=ch, left=req.steps, bytes=self.footprint(req), n=0)

    def _queue(self, rec):
        n = min(self.chunk, rec["left"])
        rec["ch"].advance(n, self.use_graph)
        rec["n"] = n

    @staticmethod
    def _wait(rec):
        """waits for rec's queued chunk; (divmax, cflm, error message or "")"""
        from .c3d import C3DDiverged
        try:
            divmax, cflm = rec["ch"].sync()
        except C3DDiverged as e:
            return e.divmax, e.cflm, str(e)
        if not (math.isfinite(divmax) and math.isfinite(cflm)):
            return divmax, cflm, f"diverged: DIVMAX = {divmax}, CFLM = {cflm}"
        return divmax, cflm, ""

    def run(self, requests, t0=None):
        """Serve `requests` (t_arrive relative to the start) to completion;
        returns the finished requests in completion order.  Requests whose
        instance diverged are in `self.failed` instead."""
        pending = sorted(requests, key=lambda r: r.t_arrive)
        for r in pending:
            if self.footprint(r) > self.budget:
                raise ValueError(f"request {r.rid} needs {self.footprint(r)} B, more than the budget")
        srpt = self.policy == "srpt"
        # resident records; `inflight` holds those with a chunk queued, in
        # queue order (fifo: every resident record)
        waiting, resident, inflight, done = [], [], deque(), []
        self.failed = []
        used = 0
        start = self.clock() if t0 is None else t0

        def fill():
            # srpt: chunks for the resident records of least remaining work
            idle = [rec for rec in resident if not rec["queued"]]
            idle.sort(key=lambda rec: rec["left"] * rec["req"].pts)
            for rec in idle:
                if len(inflight) >= self.max_active:
                    break
                self._queue(rec)
                rec["queued"] = True
                inflight.append(rec)

        def admit_ready():
            nonlocal used
            now = self.clock() - start
            while pending and pending[0].t_arrive <= now:
                waiting.append(pending.pop(0))
            if srpt:
                waiting.sort(key=lambda r: r.pts * r.steps)  # stable: ties in arrival order
            # admission in arrival (fifo) or work (srpt) order while memory
            # and slots allow
            while waiting and len(resident) < self.max_live and used + self.footprint(waiting[0]) <= self.budget:
                rec = self._admit(waiting.pop(0), self.clock() - start)
                rec["queued"] = False
                resident.append(rec)
                used += rec["bytes"]
                self.peak_bytes = max(self.peak_bytes, used)
                if not srpt:
                    self._queue(rec)
                    rec["queued"] = True
                    inflight.append(rec)
            if srpt:
                fill()

        while pending or waiting or resident:
            admit_ready()
            if not inflight:
                if pending:
                    self.sleep(max(0.0, pending[0].t_arrive - (self.clock() - start)))
                continue
            rec = inflight.popleft()
            rec["queued"] = False
            divmax, cflm, err = self._wait(rec)
            rec["left"] -= rec["n"]
            rec["divmax"], rec["cflm"] = divmax, cflm
            req = rec["req"]
            if not err and rec["left"] > 0:
                if not srpt:
                    self._queue(rec)
                    rec["queued"] = True
                    inflight.append(rec)
                continue
            req.t_done = self.clock() - start
            req.result = dict(divmax=divmax, cflm=cflm)
            if err:
                req.error = err
                self.failed.append(req)
            else:
                if self.keep_results:
                    st = rec["ch"].download()
                    req.result.update(U=st["U"], P=st["P"], PRESG=st["PRESG"])
                done.append(req)
            rec["ch"].close()
            resident.remove(rec)
            used -= rec["bytes"]
        return done


def stats(done, wall_s):
    """p50/p99 latency (arrival to completion), throughput and the HBM
    fraction of the compulsory state traffic (112 B per cell and step)."""
    lat = np.array([r.t_done - r.t_arrive for r in done])
    work = sum(r.pts * r.steps for r in done)
    thr = work / wall_s / 1e6
    return dict(requests=len(done), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), throughput_mpts_steps_s=thr,
                hbm_state_frac=112.0 * thr * 1e6 / 8e12, wall_s=wall_s)


def poisson_requests(n, rate_per_s, grids, re_taus, steps_range, seed=4, precision=1, les_precision=None):
    """A seeded Poisson arrival stream: grids and Re_tau drawn uniformly;
    Re_tau >= 1000 requests are wall-modelled LES (in `les_precision`,
    default `precision`)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(n):
        t += rng.exponential(1.0 / rate_per_s)
        g = tuple(grids[rng.integers(len(grids))])
        rt = re_taus[rng.integers(len(re_taus))]
        les = rt >= 1000
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        out.append(Request(rid=q, grid=g, re=re_from_re_tau(rt), steps=steps, seed=q % 3,
                           precision=(precision if les_precision is None else les_precision) if les else precision,
                           les=les, t_arrive=t))
    return out
