"""Continuous-batching request loop over one context.

Requests are relaxation solves or diffusion-step solves of independent
systems.  The loop admits arrived requests in arrival order while the arena's
HBM budget allows (``Context.set_budget``; an instance that would grow the
arena past it waits), keeps every admitted system's state and neighbour lists
resident, advances all running solves one residual evaluation per tick (one
launch for all of them), and retires finished solves at once, freeing their
arena slots for the next arrivals.  A request's result does not depend on what
else is in flight: every instance is evaluated with per-instance reductions in
a fixed order.

``serve_plain`` is the per-instance path for comparison: one request at a time,
in arrival order, FP64 throughout.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import DmdError, diffuse_opts, relax_opts

ERR_BUDGET = 10


@dataclass
class Request:
    rid: int
    kind: str          # "relax" or "diffuse"
    system: dict
    arrival: float     # seconds after the start of the run (or ticks, see serve)
    opts: object = None


@dataclass
class Result:
    rid: int
    kind: str
    sites: int
    state: str
    iterations: int    # Newton steps / Picard sweeps
    evals: int         # residual evaluations
    arrival: float
    admitted: float
    done: float
    q: np.ndarray = field(repr=False, default=None)
    sigma: np.ndarray = field(repr=False, default=None)
    x: np.ndarray = field(repr=False, default=None)

    @property
    def latency(self):
        return self.done - self.arrival


def default_opts(kind, mixed):
    """Relaxations: Newton-CG to |R|_2 <= 1e-8 (forcing eta 0.05) with the
    block-Jacobi preconditioner built once per solve (rebuild interval 60 =
    max_newton);
    diffusion steps: Picard sweeps to max |dx| <= 1e-10."""
    if kind == "relax":
        return relax_opts(atol=1e-8, max_newton=60, max_cg=100, eta=0.05, mixed=mixed,
                          precond=60)
    return diffuse_opts(dt=1e-9, D0=1e-7, Eb=0.4, tol=1e-10, max_iter=60, mixed=mixed)


def _start(ctx, iid, req, mixed):
    opts = req.opts if req.opts is not None else default_opts(req.kind, mixed)
    if req.kind == "relax":
        ctx.relax_start(iid, opts)
    elif req.kind == "diffuse":
        ctx.diffuse_start(iid, opts)
    else:
        raise ValueError(f"unknown request kind {req.kind!r}")


def _harvest(ctx, iid, req, admitted, now):
    st = ctx.relax_status(iid)
    q, sg, x = ctx.instance_state(iid)
    return Result(req.rid, req.kind, int(req.system["n"]), st["state"], st["newton"],
                  st["evals"], req.arrival, admitted, now, q, sg, x)


def serve(ctx, requests, budget_bytes=0, mixed=True, clock="wall", max_ticks=10 ** 7):
    """Continuous batching.  clock="wall": arrivals are seconds after the
    start and the loop sleeps while idle; clock="ticks": arrivals are tick
    numbers (deterministic admission order, for tests).  Returns the results
    in request order and the makespan."""
    ctx.set_budget(budget_bytes)
    pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
    running = {}   # iid -> (request, admitted time)
    results = {}
    t0 = time.perf_counter()
    tick = 0

    def now():
        return time.perf_counter() - t0 if clock == "wall" else float(tick)

    while (pending or running) and tick < max_ticks:
        t = now()
        # admission, in arrival order, while the budget allows
        while pending and pending[0].arrival <= t:
            req = pending[0]
            try:
                iid = ctx.add_instance(req.system)
            except DmdError as e:
                if e.code != ERR_BUDGET:
                    raise
                if not running:
                    raise RuntimeError(f"request {req.rid} ({req.system['n']} sites) does not "
                                       "fit the budget even alone") from e
                break
            _start(ctx, iid, req, mixed)
            running[iid] = (req, t)
            pending.pop(0)
        if not running:
            if clock == "wall":
                time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
            else:
                tick = int(np.ceil(pending[0].arrival))
            continue
        ctx.relax_tick()
        tick += 1
        t = now()
        for iid in list(running):
            if ctx.relax_status(iid)["state"] != "running":
                req, adm = running.pop(iid)
                results[req.rid] = _harvest(ctx, iid, req, adm, t)
                ctx.remove_instance(iid)
    ctx.set_budget(0)
    return [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)], now()


def serve_plain(ctx, requests, clock="wall"):
    """One request at a time in arrival order, FP64 residuals and probes."""
    order = sorted(requests, key=lambda r: (r.arrival, r.rid))
    results = {}
    t0 = time.perf_counter()
    tick = 0

    def now():
        return time.perf_counter() - t0 if clock == "wall" else float(tick)

    for req in order:
        if now() < req.arrival:
            if clock == "wall":
                time.sleep(req.arrival - now())
            else:
                tick = int(np.ceil(req.arrival))
        adm = now()
        iid = ctx.add_instance(req.system)
        _start(ctx, iid, req, False)
        while ctx.relax_tick():
            tick += 1
        tick += 1
        results[req.rid] = _harvest(ctx, iid, req, adm, now())
        ctx.remove_instance(iid)
    return [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)], now()


def latency_stats(results):
    lat = np.array([r.latency for r in results])
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
            "mean": float(lat.mean()), "max": float(lat.max())}


def poisson_arrivals(n, rate, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n)) - 0.0
