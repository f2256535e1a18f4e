"""Host logic without a GPU: the dump reader on the reference's own input
file, the dump writer's round trip, the seeded workload generators the
benchmarks use, and the continuous-batching loop's admission, budget and
retirement rules driven through a stand-in context.

The dump fixture is the reference's inputs/Mg-hcp-cube-x5-x5-x5.dump (500 Mg
atoms, a = 3.189 A, the ideal c/a = sqrt(8/3) = 1.633, sigma = 0.01,
beta = 1/(k_B 300 K)).  The generator counts are those of the benchmark
workloads (the 54,000-site box, the 64-instance nanowire/sphere batch, the
200-request serving mix): a change to a generator changes the benchmarks, and
these tests say so.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, RC

import dmdgpu
import dmdgpu.lattice as lat
from dmdgpu import serve


def test_reference_dump_reads_as_its_lattice():
    sd = lat.read_dump(os.path.join(GOLDEN, "Mg-hcp-cube-x5-x5-x5.dump"))
    assert sd["n"] == 500
    assert np.all(sd["species"] == lat.MG)
    assert sd["periodic"].tolist() == [1, 1, 1]
    a, c = 3.189, 3.189 * np.sqrt(8.0 / 3.0)
    L = sd["hi"] - sd["lo"]
    assert np.allclose(L, [5 * a, 5 * np.sqrt(3.0) * a, 5 * c], rtol=0, atol=1e-6)
    assert np.all(sd["sigma"] == 0.01) and np.all(sd["x"] == 1.0) and np.all(sd["gamma"] == 0.0)
    assert np.allclose(sd["beta"], 1.0 / (lat.K_B * 300.0), rtol=1e-10, atol=0)
    # 4 atoms per orthorhombic cell on the hcp sites: fractional coordinates
    # of every atom are multiples of the basis offsets
    f = (sd["q"] - sd["lo"]) / (L / 5.0)
    frac = np.mod(f, 1.0)
    basis = lat.HCP_BASIS
    d = np.abs(frac[:, None, :] - basis[None, :, :])
    d = np.minimum(d, 1.0 - d).max(axis=2).min(axis=1)
    assert d.max() < 1e-6


@pytest.mark.parametrize("make", [lambda: lat.read_dump(os.path.join(GOLDEN, "Mg-hcp-cube-x5-x5-x5.dump")),
                                  lambda: lat.nanowire(12.0, seed=3),
                                  lambda: lat.cluster(3, 2, 2, x_h=(0.0, 0.4), seed=5)])
def test_dump_round_trip_is_bit_exact(tmp_path, make):
    sd = make()
    path = str(tmp_path / "sys.dump")
    lat.write_dump(path, sd, step=7)
    back = lat.read_dump(path)
    for key in ("q", "sigma", "x", "species", "beta", "gamma", "lo", "hi", "periodic"):
        assert np.array_equal(back[key], sd[key]), key
    assert back["n"] == sd["n"]


def test_dump_reader_rejects_other_species_and_flags(tmp_path):
    sd = lat.cluster(2, 1, 1, seed=1)
    path = str(tmp_path / "bad.dump")
    lat.write_dump(path, sd)
    text = open(path).read().splitlines()
    bad = text[:9] + ["13" + text[9][text[9].index(" "):]] + text[10:]
    open(path, "w").write("\n".join(bad) + "\n")
    with pytest.raises(ValueError):
        lat.read_dump(path)
    lat.write_dump(path, sd)
    text = open(path).read().splitlines()
    text[4] = "ITEM: BOX BOUNDS pp fm pp"
    open(path, "w").write("\n".join(text) + "\n")
    with pytest.raises(ValueError):
        lat.read_dump(path)


def test_headline_workload():
    sd = lat.hcp_mgh(20, 15, 15, x_h=(0.0, 0.1), seed=2024)
    assert sd["n"] == 54_000
    assert int((sd["species"] == lat.MG).sum()) == 18_000
    assert int((sd["species"] == lat.H).sum()) == 36_000
    assert sd["periodic"].tolist() == [1, 1, 1]
    h = sd["species"] == lat.H
    assert sd["x"][h].min() >= 0.0 and sd["x"][h].max() <= 0.1
    assert np.all(sd["x"][~h] == 1.0)
    # same seed, same system
    again = lat.hcp_mgh(20, 15, 15, x_h=(0.0, 0.1), seed=2024)
    assert np.array_equal(again["q"], sd["q"]) and np.array_equal(again["x"], sd["x"])


def test_batched_relaxation_workload():
    batch = lat.mixed_batch(64, seed=2024)
    ns = [s["n"] for s in batch]
    assert (min(ns), max(ns), sum(ns)) == (949, 19_020, 424_184)
    assert all(s["periodic"].tolist() == [0, 0, 1] for s in batch[0::2])  # nanowires
    assert all(s["periodic"].tolist() == [0, 0, 0] for s in batch[1::2])  # spheres
    for s in batch:
        inside = (s["q"] >= s["lo"] - 1e-9) & (s["q"] <= s["hi"] + 1e-9)
        assert inside[:, 2].all()  # the periodic axis is wrapped into the box


def test_serving_mix_and_box_sizes():
    mix = lat.serving_mix(200, seed=7)
    ns = [s["n"] for _, s in mix]
    assert (min(ns), max(ns), sum(ns)) == (540, 199_584, 5_482_422)
    assert sum(k == "relax" for k, _ in mix) == 102
    for n in (600, 2000, 6000, 30000, 155000):
        sd = lat.box_of_size(n, seed=1)
        assert sd["periodic"].tolist() == [1, 1, 1]
        assert np.all(sd["hi"] - sd["lo"] >= 2.0 * RC)  # every periodic edge >= 2 r_c
        assert 0.85 * n <= sd["n"] <= 1.05 * n or n < 1000


class FakeContext:
    """The calls serve() and serve_plain() make on a dmdgpu.Context: an
    instance of n sites holds n * per_site bytes and its solve finishes after
    ticks(n) ticks; add_instance refuses with DMDGPU_ERR_BUDGET past the
    budget, as the library does."""

    def __init__(self, per_site=100, ticks=lambda n: 1 + n // 100):
        self.per_site, self.ticks = per_site, ticks
        self.budget, self.inst, self.next, self.peak = 0, {}, 0, 0
        self.tick_count, self.live_per_tick = 0, []

    def set_budget(self, b):
        self.budget = b

    def add_instance(self, sysd):
        held = (sum(v["n"] for v in self.inst.values()) + sysd["n"]) * self.per_site
        if self.budget and held > self.budget:
            raise dmdgpu.DmdError(serve.ERR_BUDGET, "instance does not fit the budget")
        iid, self.next = self.next, self.next + 1
        self.inst[iid] = {"n": sysd["n"], "state": "idle", "left": 0, "done": 0, "sys": sysd}
        self.peak = max(self.peak, held)
        return iid

    def _start(self, iid):
        v = self.inst[iid]
        v["state"], v["left"] = "running", self.ticks(v["n"])

    def relax_start(self, iid, opts):
        self._start(iid)

    def diffuse_start(self, iid, opts):
        self._start(iid)

    def relax_tick(self):
        live = [v for v in self.inst.values() if v["state"] == "running"]
        self.tick_count += 1
        self.live_per_tick.append(len(live))
        for v in live:
            v["left"] -= 1
            v["done"] += 1
            if v["left"] == 0:
                v["state"] = "converged"
        return any(v["state"] == "running" for v in self.inst.values())

    def relax_status(self, iid):
        v = self.inst[iid]
        return {"state": v["state"], "newton": v["done"], "evals": v["done"]}

    def instance_state(self, iid):
        s = self.inst[iid]["sys"]
        return s["q"], s["sigma"], s["x"]

    def remove_instance(self, iid):
        del self.inst[iid]


def _requests(sizes, arrivals, kinds=None):
    kinds = kinds or ["relax"] * len(sizes)
    return [serve.Request(i, k, {"n": n, "q": np.zeros((n, 3)), "sigma": np.ones(n),
                                 "x": np.ones(n)}, float(a))
            for i, (n, a, k) in enumerate(zip(sizes, arrivals, kinds))]


def test_serving_admits_in_arrival_order_within_the_budget():
    sizes = [300, 900, 200, 700, 100, 500, 400]
    arrivals = [0, 0, 1, 2, 2, 3, 30]
    kinds = ["relax", "diffuse"] * 3 + ["relax"]
    reqs = _requests(sizes, arrivals, kinds)
    ctx = FakeContext()
    budget = 1200 * ctx.per_site
    res, makespan = serve.serve(ctx, reqs, budget_bytes=budget, clock="ticks")
    assert [r.rid for r in res] == list(range(len(reqs)))
    assert all(r.state == "converged" for r in res)
    assert ctx.peak <= budget and ctx.budget == 0  # budget restored at the end
    assert not ctx.inst  # every finished instance was removed
    for r in res:
        assert r.admitted >= r.arrival
        assert r.iterations == ctx.ticks(r.sites)
        assert r.done - r.admitted == r.iterations  # one evaluation per tick
        assert r.kind == kinds[r.rid]
    # first come, first served: admission follows arrival, so a request that
    # does not fit holds the ones behind it
    order = sorted(res, key=lambda r: (r.arrival, r.rid))
    assert all(a.admitted <= b.admitted for a, b in zip(order, order[1:]))
    # several solves share ticks, and the idle gap before the last arrival is skipped
    assert max(ctx.live_per_tick) >= 2
    assert res[-1].admitted == 30.0 and makespan == res[-1].done


def test_serving_request_larger_than_the_budget_alone_fails():
    reqs = _requests([100, 5000], [0, 0])
    ctx = FakeContext()
    with pytest.raises(RuntimeError):
        # the first runs; the second waits for it, then cannot fit even alone
        serve.serve(ctx, reqs, budget_bytes=1000 * ctx.per_site, clock="ticks")


def test_plain_server_runs_one_request_at_a_time():
    sizes, arrivals = [300, 200, 500, 100], [0, 0, 1, 50]
    reqs = _requests(sizes, arrivals)
    ctx = FakeContext()
    res, makespan = serve.serve_plain(ctx, reqs, clock="ticks")
    assert [r.rid for r in res] == [0, 1, 2, 3]
    assert max(ctx.live_per_tick) == 1
    order = sorted(res, key=lambda r: (r.arrival, r.rid))
    assert all(b.admitted >= a.done for a, b in zip(order, order[1:]))
    assert res[3].admitted == 50.0 and makespan == res[3].done


def test_latency_statistics_and_arrivals():
    R = [serve.Result(i, "relax", 10, "converged", 1, 1, 0.0, 0.0, float(i + 1)) for i in range(100)]
    st = serve.latency_stats(R)
    assert st["p50"] == pytest.approx(50.5) and st["p99"] == pytest.approx(99.01)
    assert st["max"] == 100.0 and st["mean"] == pytest.approx(50.5)
    t = serve.poisson_arrivals(20000, rate=25.0, seed=3)
    assert np.all(np.diff(t) > 0)
    assert t[-1] / len(t) == pytest.approx(1 / 25.0, rel=0.03)
