"""GPU tests of the diffusion (occupancy) step against the oracle.

* diffusive neighbour lists equal the oracle's brute-force lists;
* the rates xdot at a state match the oracle (mu from the oracle residual)
  to 1e-10 relative;
* one implicit step (Picard sweeps, FP64 residuals) matches the oracle's step
  to 1e-10, conserves the hydrogen content, and the split-precision step stays
  within 1e-8 of it;
* a diffusion step run in a mix with relaxations is bitwise the step alone.
"""
import sys

import numpy as np
import pytest

from conftest import RC

import dmdgpu
import dmdgpu.lattice as lat

pytestmark = pytest.mark.gpu

D0, EB, DT = 1e-7, 0.4, 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = dmdgpu.Context(0)
    yield c
    c.close()


def systems():
    return [lat.hcp_mgh(4, 3, 3, seed=1, x_h=(0.0, 0.3)), lat.nanowire(15, seed=3)]


def fresh(ctx, sd):
    ctx.set_state(sd)
    ctx.neighbors()
    return 0


def test_diffusive_lists(ctx, oracle):
    for sd in systems():
        say(f"[gpu] diffusive lists n={sd['n']}")
        iid = fresh(ctx, sd)
        cnt, lst = ctx.diffusion_neighbors(iid)
        st, ocnt, olst = oracle.diff_neighbors(sd)
        assert st == 0
        assert np.array_equal(cnt, ocnt)
        for i in range(sd["n"]):
            assert sorted(lst[i, :cnt[i]].tolist()) == sorted(olst[i, :ocnt[i]].tolist())


def test_rates_match_oracle(ctx, oracle):
    for sd in systems():
        say(f"[gpu] diffusion rates n={sd['n']}")
        iid = fresh(ctx, sd)
        xd = ctx.diffusion_rates(iid, dmdgpu.diffuse_opts(D0=D0, Eb=EB))
        st, nnb, nbr = oracle.neighbors(sd, RC)
        res = oracle.fact_residual(oracle.system(sd, nnb, nbr))
        mu = oracle.diff_mu(sd, res["Rx"])
        st, cnt, lst = oracle.diff_neighbors(sd)
        ref = oracle.diff_rates(sd, mu, cnt, lst, D0, EB)
        err = np.abs(xd - ref).max() / np.abs(ref).max()
        say(f"    max rel err {err:.2e}")
        assert err < 1e-10
        assert abs(xd.sum()) <= 1e-12 * np.abs(xd).sum()


@pytest.mark.parametrize("mixed", [False, True])
def test_step_matches_oracle(ctx, oracle, mixed):
    sd = systems()[1]
    say(f"[gpu] diffusion step mixed={mixed} n={sd['n']}")
    iid = fresh(ctx, sd)
    ctx.diffuse_start(iid, dmdgpu.diffuse_opts(dt=DT, D0=D0, Eb=EB, tol=1e-10, mixed=mixed))
    ticks = 0
    while ctx.relax_tick() and ticks < 200:
        ticks += 1
    stat = ctx.relax_status(iid)
    say(f"    {stat}")
    assert stat["state"] == "converged"
    _, _, x = ctx.instance_state(iid)
    st, nnb, nbr = oracle.neighbors(sd, RC)
    xo, sweeps, dmax = oracle.diffusion_step(sd, nnb, nbr, dt=DT, D0=D0, Eb=EB, tol=1e-10)
    h = sd["species"] == 1
    err = np.abs(x - xo).max()
    say(f"    oracle sweeps {sweeps}, gpu sweeps {stat['newton']}, max |x - x_oracle| {err:.2e}")
    assert err < (1e-8 if mixed else 1e-10)
    assert abs(x[h].sum() - sd["x"][h].sum()) <= 1e-12 * sd["x"][h].sum()
    assert np.abs(x - sd["x"]).max() > 1e-4


def test_step_in_mix_equals_alone(ctx):
    say("[gpu] diffusion step in a mix with relaxations")
    dsys = lat.hcp_mgh(4, 3, 3, seed=21, x_h=(0.0, 0.3))
    rsys = [lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=s) for s in (31, 32)]
    dop = dmdgpu.diffuse_opts(dt=DT, D0=D0, Eb=EB, tol=1e-10, mixed=True)
    rop = dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=0.1, mixed=True)
    iid = fresh(ctx, dsys)
    ctx.diffuse_start(iid, dop)
    while ctx.relax_tick():
        pass
    alone = ctx.instance_state(iid)[2], ctx.relax_status(iid)
    ctx.remove_instance(iid)
    r0 = ctx.add_instance(rsys[0])
    ctx.relax_start(r0, rop)
    for _ in range(5):
        ctx.relax_tick()
    d = ctx.add_instance(dsys)
    ctx.diffuse_start(d, dop)
    r1 = ctx.add_instance(rsys[1])
    ctx.relax_start(r1, rop)
    while ctx.relax_tick():
        pass
    x = ctx.instance_state(d)[2]
    assert ctx.relax_status(d) == alone[1]
    assert np.array_equal(x, alone[0])
    for i in (r0, r1):
        assert ctx.relax_status(i)["state"] == "converged"
    for i in (d, r0, r1):
        ctx.remove_instance(i)
