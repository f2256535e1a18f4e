"""GPU test of the continuous-batching request loop.

Relaxation and diffusion-step requests arrive while others are on the GPU;
the HBM budget admits only some of them at a time, so later requests wait for
slots and memory to free up.  Each result must be bitwise the result of the
same request served alone (the solves are deterministic per instance), which
is tighter than the stated tolerances (|R| <= 1e-8 for relaxations,
max |dx| <= 1e-10 for diffusion steps).
"""
import sys

import numpy as np
import pytest

import dmdgpu
import dmdgpu.lattice as lat
from dmdgpu.serve import Request, latency_stats, serve, serve_plain

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = dmdgpu.Context(0)
    yield c
    c.close()


def requests():
    syss = [("relax", lat.box_of_size(500, seed=1)), ("diffuse", lat.box_of_size(900, seed=2)),
            ("relax", lat.wire_of_size(1500, seed=3)), ("diffuse", lat.wire_of_size(800, seed=4)),
            ("relax", lat.box_of_size(1200, seed=5)), ("diffuse", lat.box_of_size(600, seed=6))]
    arrivals = [0, 0, 3, 5, 20, 21]  # ticks
    return [Request(k, kind, sd, float(a)) for k, ((kind, sd), a) in
            enumerate(zip(syss, arrivals))]


def test_results_same_alone_and_in_mix(ctx):
    reqs = requests()
    say(f"[gpu] serving {len(reqs)} requests, sizes {[r.system['n'] for r in reqs]}")
    per_site = ctx.bytes_per_site()
    budget = per_site * 2 * 2048  # about two of these instances at a time
    mixed, span = serve(ctx, reqs, budget_bytes=budget, mixed=True, clock="ticks")
    say(f"    served in {span:.0f} ticks; latency (ticks) {latency_stats(mixed)}")
    # some request had to wait for memory
    assert any(r.admitted > r.arrival for r in mixed)
    for req, res in zip(reqs, mixed):
        alone, _ = serve(ctx, [Request(req.rid, req.kind, req.system, 0.0)], mixed=True,
                         clock="ticks")
        a = alone[0]
        say(f"    {req.kind} n={res.sites}: {res.state} it={res.iterations} evals={res.evals}")
        assert res.state == "converged"
        assert (res.state, res.iterations, res.evals) == (a.state, a.iterations, a.evals)
        assert np.array_equal(res.q, a.q)
        assert np.array_equal(res.sigma, a.sigma)
        assert np.array_equal(res.x, a.x)


def test_plain_server_converges(ctx):
    reqs = requests()[:3]
    res, span = serve_plain(ctx, reqs, clock="ticks")
    for r in res:
        assert r.state == "converged"
    # plain and batched solve the same equations: same converged states to tolerance
    bat, _ = serve(ctx, reqs, mixed=True, clock="ticks")
    for p, b in zip(res, bat):
        assert np.abs(p.q - b.q).max() < 1e-6
        assert np.abs(p.x - b.x).max() < 1e-8
