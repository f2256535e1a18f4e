"""GPU tests of resident instances (batching) and the Newton-Krylov relaxation.

* Several instances evaluated in one launch give, per instance, bitwise the
  same residual as that instance evaluated alone, and match the oracle.
* Relaxation drives |(R_q, R_sigma)|_2 below atol; the oracle confirms the
  final state's residual.  The mixed mode (split-precision Hessian probes,
  FP64 residuals) reaches the same tolerance.
* A relaxation inside a batch follows bitwise the same trajectory as alone.
* Relaxation ticks skip the occupancy residual R_x (and, with the residual
  merit, the energies); the residual read after a relaxation still carries the
  full R_x and totals of the final state, under every pair rule and in both
  precisions.
"""
import sys

import numpy as np
import pytest

from conftest import RC

import dmdgpu
import dmdgpu.lattice as lat

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def systems():
    return [lat.cluster(3, 2, 2, x_h=(0.0, 0.3), seed=41),
            lat.hcp_mgh(5, 3, 3, x_h=(0.0, 0.2), seed=42),
            lat.hcp_mgh(4, 3, 3, with_h=False, seed=43),
            lat.cluster(2, 2, 1, x_h=(0.0, 0.5), seed=44)]


@pytest.fixture(scope="module")
def ctx():
    c = dmdgpu.Context(0)
    yield c
    c.close()


def oracle_residual(oracle, sysd):
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    assert st == 0
    return oracle.fact_residual(oracle.system(sysd, nnb, nbr))


def test_batch_equals_alone_and_oracle(ctx, oracle):
    say("[gpu] batch vs alone")
    syss = systems()
    alone = []
    for sd in syss:
        ctx.set_state(sd)
        ctx.neighbors()
        alone.append(ctx.residual())
    ctx.set_state(syss[0])  # clears the arena
    ctx.remove_instance(0)
    ids = [ctx.add_instance(sd) for sd in syss]
    ctx.evaluate()
    for k, (iid, sd) in enumerate(zip(ids, syss)):
        out = ctx.instance_residual(iid)
        for key in ("Rq", "Rs", "Rx"):
            assert np.array_equal(out[key], alone[k][key]), (k, key)
        assert out["S0_total"] == alone[k]["S0_total"]
        ref = oracle_residual(oracle, sd)
        assert rel(out["Rq"], ref["Rq"]) < 1e-10
        assert rel(out["Rs"], ref["Rs"]) < 1e-10


def test_slot_reuse_and_budget(ctx):
    say("[gpu] slot reuse + budget")
    syss = systems()
    ctx.set_state(syss[0])
    ctx.remove_instance(0)
    a = ctx.add_instance(syss[1])
    b = ctx.add_instance(syss[2])
    ctx.evaluate()
    ref_b = ctx.instance_residual(b)
    ctx.remove_instance(a)
    c = ctx.add_instance(syss[3])  # fits into a's freed slots
    ctx.evaluate()
    assert np.array_equal(ctx.instance_residual(b)["Rq"], ref_b["Rq"])
    info = ctx.instance_info(c)
    assert info["n"] == syss[3]["n"]
    ctx.set_budget(ctx.bytes_per_site() * info["arena_sites"])
    with pytest.raises(dmdgpu.DmdError) as e:
        ctx.add_instance(syss[1])  # would grow the arena past the budget
    assert e.value.code == 10
    ctx.set_budget(0)
    ctx.remove_instance(b)
    ctx.remove_instance(c)


@pytest.mark.parametrize("mixed,tuning", [(False, None), (True, None),
                                          (True, ((8, 64, 0), (8, 64, 1)))])
def test_relaxation_converges(ctx, oracle, mixed, tuning):
    """tuning: (FP64, split) kernel tuning, None for the context default; the
    third case is the tuning the benchmark's autotuner selects."""
    say(f"[gpu] relax mixed={mixed} tuning={tuning}")
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=7)
    if tuning:
        ctx.set_tuning_precision("fp64", *tuning[0])
        ctx.set_tuning_precision("split32", *tuning[1])
    else:
        ctx.set_tuning(4, 4, 1)
    ctx.set_state(sd)
    ctx.neighbors()
    r0 = ctx.residual()
    n0 = np.sqrt((r0["Rq"] ** 2).sum() + (r0["Rs"] ** 2).sum())
    opts = dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=1e-3, mixed=mixed)
    ctx.relax_start(0, opts)
    ticks = 0
    while ctx.relax_tick() and ticks < 20000:
        ticks += 1
    st = ctx.relax_status(0)
    say(f"    {st} ticks={ticks} |R0|={n0:.3e}")
    assert st["state"] == "converged", st
    q, sg, x = ctx.instance_state(0)
    fin = dict(sd)
    fin["q"], fin["sigma"] = q, sg
    ref = oracle_residual(oracle, fin)
    # default unknowns: Mg positions and widths (H sites stay at their sites)
    mg = (sd["species"] == 12) & (sd["x"] >= 5e-4)
    rn = np.sqrt((ref["Rq"][mg] ** 2).sum() + (ref["Rs"][mg] ** 2).sum())
    assert rn < 1e-7, rn
    assert np.abs(q - sd["q"]).max() > 1e-4  # it moved
    # Relaxation ticks leave the occupancy residual out; asking for the
    # instance's residual afterwards evaluates it at the final state.
    out = ctx.instance_residual(0)
    assert rel(out["Rx"], ref["Rx"]) < 1e-10
    assert rel(out["Rq"], ref["Rq"]) < 1e-10
    ctx.set_tuning(4, 4, 1)


MG_QS = dmdgpu.DOF_MG_Q | dmdgpu.DOF_MG_SIGMA


@pytest.mark.parametrize("dof,rule,prec", [(MG_QS, 3, "fp64"),
                                           (MG_QS | dmdgpu.DOF_H_SIGMA, 3, "fp64"),
                                           (MG_QS, 3, "split32"), (MG_QS, 5, "fp64"),
                                           (MG_QS, 44, "split32"), (MG_QS, 6, "fp64")])
def test_relaxation_tick_residual_equals_full_evaluation(ctx, dof, rule, prec):
    """The first tick of a relaxation evaluates the residual at the start
    point the way relaxation ticks do: without R_x, without energies (the
    pair-rule points give densities only) and, when only Mg unknowns are
    relaxed, without the pairs with two H ends and the H sites' sums; FP64
    probes start in FP64, split-precision probes in split precision.  The norm
    it reports over the relaxed unknowns equals the norm of the same rows of a
    full evaluation in that precision (to the summation order's rounding),
    under every pair rule, and reading the residual afterwards gives the full
    evaluation bit for bit."""
    say(f"[gpu] relaxation tick residual vs full evaluation, dof {dof}, rule {rule}, {prec}")
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=7)
    ctx.set_state(sd)
    ctx.neighbors()
    ctx.set_rule(rule)
    ctx.set_precision(prec)
    try:
        _tick_vs_full(ctx, sd, dof, prec == "split32")
    finally:
        ctx.set_rule(3)
        ctx.set_precision("fp64")


def _tick_vs_full(ctx, sd, dof, mixed):
    full = ctx.residual()
    act = sd["x"] >= 5e-4
    mg = (sd["species"] == lat.MG) & act
    h = (sd["species"] == lat.H) & act
    n2 = (full["Rq"][mg] ** 2).sum() + (full["Rs"][mg] ** 2).sum()
    if dof & dmdgpu.DOF_H_SIGMA:
        n2 += (full["Rs"][h] ** 2).sum()
    ctx.relax_start(0, dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=0.05,
                                         mixed=mixed, dof_mask=dof))
    ctx.relax_tick()
    st = ctx.relax_status(0)
    say(f"    tick norm {st['rnorm']:.15e}, full {np.sqrt(n2):.15e}")
    assert st["rnorm"] == pytest.approx(np.sqrt(n2), rel=1e-13)
    again = ctx.instance_residual(0)
    for k in ("Rq", "Rs", "Rx"):
        assert np.array_equal(again[k], full[k]), k


@pytest.mark.parametrize("mixed,every", [(False, 1), (True, 1), (True, 40)])
def test_preconditioned_relaxation(ctx, oracle, mixed, every):
    """Block-Jacobi preconditioned CG (diagonal Hessian blocks inverted per
    site, rebuilt every `every` Newton steps): converges to the same bar in
    fewer residual evaluations than plain CG, to the same state."""
    say(f"[gpu] preconditioned relax mixed={mixed} every={every}")
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=7)
    res = {}
    for pc in (0, every):
        ctx.set_state(sd)
        ctx.neighbors()
        opts = dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=0.05, mixed=mixed,
                                 precond=pc)
        ctx.relax_start(0, opts)
        ticks = 0
        while ctx.relax_tick() and ticks < 20000:
            ticks += 1
        st = ctx.relax_status(0)
        say(f"    precond={pc}: {st} ticks={ticks}")
        assert st["state"] == "converged", st
        res[pc] = (ctx.instance_state(0), st)
    (q, sg, _), st = res[every]
    fin = dict(sd)
    fin["q"], fin["sigma"] = q, sg
    ref = oracle_residual(oracle, fin)
    mg = (sd["species"] == 12) & (sd["x"] >= 5e-4)
    rn = np.sqrt((ref["Rq"][mg] ** 2).sum() + (ref["Rs"][mg] ** 2).sum())
    assert rn < 1e-7, rn
    assert np.abs(q - res[0][0][0]).max() < 1e-7
    assert st["evals"] < res[0][1]["evals"]


@pytest.mark.parametrize("dof,merit", [(7, "residual"), (7, "energy"), (15, "residual")])
def test_relaxation_h_positions_bulk(ctx, oracle, dof, merit):
    """H positions relaxed with the Mg lattice (dof 7 = Mg q, Mg sigma, H q;
    15 = every unknown) in a periodic Mg-H box: the preconditioned Newton-CG
    converges with either line-search merit, confirmed by the oracle's
    residual over the relaxed unknowns."""
    say(f"[gpu] relax with H positions, periodic box, dof {dof}, merit {merit}")
    sd = lat.hcp_mgh(4, 3, 3, x_h=(0.05, 0.3), seed=7)
    ctx.set_state(sd)
    ctx.neighbors()
    opts = dmdgpu.relax_opts(atol=1e-8, max_newton=60, max_cg=100, eta=0.05, precond=60,
                             dof_mask=dof, merit=merit)
    ctx.relax(opts)
    st = ctx.relax_status(0)
    say(f"    {st}")
    assert st["state"] == "converged", st
    q, sg, _ = ctx.instance_state(0)
    fin = dict(sd)
    fin["q"], fin["sigma"] = q, sg
    # the solve keeps the neighbour lists of its start state (as the
    # reference's SNES keeps its topology): the oracle uses the same lists
    st0, nnb, nbr = oracle.neighbors(sd, RC)
    ref = oracle.fact_residual(oracle.system(fin, nnb, nbr))
    act = sd["x"] >= 5e-4
    mg = (sd["species"] == 12) & act
    h = (sd["species"] == 1) & act
    rs = mg | h if dof & dmdgpu.DOF_H_SIGMA else mg
    rn = np.sqrt((ref["Rq"][mg | h] ** 2).sum() + (ref["Rs"][rs] ** 2).sum())
    say(f"    oracle |R| {rn:.3e}, max H displacement {np.abs(q - sd['q'])[h].max():.3f} A")
    assert rn < 1e-7, rn
    assert np.abs(q - sd["q"])[h].max() > 1e-3  # the H sites moved


def test_energy_merit_descends_where_residual_merit_stalls(ctx):
    """A free 144-site Mg-H cluster with H positions relaxed reconstructs far
    from its start (atoms move more than 4 A before a minimum is reached).
    There the residual-norm line search stalls, while the energy merit keeps
    lowering the free energy: after the same number of evaluations it is at a
    lower energy.  (Which of the two unconverged states has the smaller
    residual depends on the rounding of a chaotic trajectory, so it is only
    printed.)"""
    say("[gpu] energy merit on a reconstructing cluster")
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=7)
    dof = dmdgpu.DOF_MG_Q | dmdgpu.DOF_MG_SIGMA | dmdgpu.DOF_H_Q
    out = {}
    for merit in ("residual", "energy"):
        ctx.set_state(sd)
        ctx.neighbors()
        ctx.relax(dmdgpu.relax_opts(atol=1e-8, max_newton=400, max_cg=100, eta=0.05,
                                    dof_mask=dof, merit=merit, max_evals=2500))
        st = ctx.relax_status(0)
        out[merit] = (st, ctx.relax_energy(0))
        say(f"    {merit}: {st}, energy {out[merit][1]:.6f}")
    e0 = None
    ctx.set_state(sd)
    ctx.neighbors()
    e0 = ctx.residual()["S0_total"]
    assert out["residual"][0]["state"] in ("stalled", "maxiter")
    assert out["energy"][1] < out["residual"][1] < e0
    say(f"    rnorm: energy merit {out['energy'][0]['rnorm']:.4g}, "
        f"residual merit {out['residual'][0]['rnorm']:.4g}")


def test_residual_merit_energy_of_final_iterate(ctx):
    """Relaxation ticks with the residual-norm line search form no energies
    (the pair-rule points give densities only); relax_energy then evaluates
    the final iterate in FP64 once the solve has ended: bit for bit the total
    of a full evaluation at that state, and the residual read after the solve
    is the full one (its totals included)."""
    say("[gpu] residual-merit ticks: energy of the final iterate")
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=31)
    ctx.set_state(sd)
    ctx.neighbors()
    ctx.relax(dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=0.05, precond=40))
    st = ctx.relax_status(0)
    assert st["state"] == "converged", st
    after = ctx.instance_residual(0)  # evaluated again: whole residual and totals
    e = ctx.relax_energy(0)
    q, sg, _ = ctx.instance_state(0)
    ctx.update_state(q=q, sigma=sg)
    full = ctx.residual()
    assert e == full["S0_total"] and after["S0_total"] == full["S0_total"]
    for k in ("Rq", "Rs", "Rx"):
        assert np.array_equal(after[k], full[k]), k


@pytest.mark.skip(reason="the free 144-site cluster with H positions relaxed reconstructs: on the "
                         "CPU prototype the energy merit needs about 440 Newton steps and 32k "
                         "evaluations to reach |R| = 3e-5 (atoms move > 4 A); the residual merit "
                         "stalls at |R| = 7.6 (|R| = 8.5 after 811 evaluations on MI355X)")
def test_preconditioned_relaxation_h_positions(ctx, oracle):
    """Mg positions, Mg widths and H positions relaxed (dof 7) on the free
    cluster, preconditioned CG rebuilt every 40 Newton steps, FP64 probes:
    converges, confirmed by the oracle's residual over the relaxed unknowns."""
    say("[gpu] preconditioned relax with H positions")
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=7)
    ctx.set_state(sd)
    ctx.neighbors()
    dof = dmdgpu.DOF_MG_Q | dmdgpu.DOF_MG_SIGMA | dmdgpu.DOF_H_Q
    opts = dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=0.05, precond=40,
                             dof_mask=dof)
    ctx.relax(opts)
    st = ctx.relax_status(0)
    say(f"    {st}")
    assert st["state"] == "converged", st


@pytest.mark.parametrize("precond", [False, True])
def test_relaxation_batch_equals_alone(ctx, precond):
    say(f"[gpu] relax batch vs alone, precond={precond}")
    syss = [lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=s) for s in (11, 12, 13)]
    opts = dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=1e-3, mixed=True,
                             precond=precond)
    alone = []
    for sd in syss:
        ctx.set_state(sd)
        ctx.neighbors()
        ctx.relax(opts)
        alone.append((ctx.instance_state(0), ctx.relax_status(0)))
    ctx.remove_instance(0)
    ids = [ctx.add_instance(sd) for sd in syss]
    ctx.relax(opts)
    for k, iid in enumerate(ids):
        q, sg, _ = ctx.instance_state(iid)
        st = ctx.relax_status(iid)
        assert st == alone[k][1], (st, alone[k][1])
        assert np.array_equal(q, alone[k][0][0])
        assert np.array_equal(sg, alone[k][0][1])
        assert st["state"] == "converged"
    for iid in ids:
        ctx.remove_instance(iid)


def test_checkpoint_resume(ctx, tmp_path):
    """A relaxation stopped after 40 evaluations, written with
    dmdgpu.checkpoint and resumed from the files, reaches the state of the
    uninterrupted relaxation."""
    import dmdgpu.checkpoint as ck
    say("[gpu] checkpoint resume")
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=7)
    opts = dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=0.05, precond=40)
    ctx.set_state(sd)
    ctx.neighbors()
    ctx.relax(opts)
    assert ctx.relax_status(0)["state"] == "converged"
    q_ref = ctx.instance_state(0)[0]
    ctx.set_state(sd)
    ctx.neighbors()
    part = dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=0.05, precond=40,
                             max_evals=40)
    ctx.relax(part)
    assert ctx.relax_status(0)["state"] == "maxiter"
    q, sg, x = ctx.instance_state(0)
    mid = dict(sd)
    mid["q"], mid["sigma"], mid["x"] = q, sg, x
    ck.write_checkpoint(str(tmp_path), mid, step=1)
    back, _ = ck.read_checkpoint(str(tmp_path))
    ctx.set_state(back)
    ctx.neighbors()
    ctx.relax(opts)
    assert ctx.relax_status(0)["state"] == "converged"
    assert np.abs(ctx.instance_state(0)[0] - q_ref).max() < 1e-7


@pytest.mark.parametrize("mixed", [False, True])
def test_graph_ticks_equal_direct_launches(mixed):
    """The tick's device work captured once as a hipGraph (FP64 and split
    chains on two streams, joined) and replayed gives bit for bit the
    relaxations of the same ticks launched directly, for a batch whose
    instances run in different precisions in the same ticks."""
    say(f"[gpu] graph ticks vs direct launches, mixed={mixed}")
    syss = [lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=s) for s in (21, 22)] + \
        [lat.hcp_mgh(4, 3, 3, with_h=False, seed=23)]
    opts = dmdgpu.relax_opts(atol=1e-8, max_newton=40, max_cg=80, eta=0.05, mixed=mixed,
                             precond=40)
    res = {}
    for graph in (False, True):
        c = dmdgpu.Context(0)
        try:
            c.set_graph(graph)
            ids = [c.add_instance(sd) for sd in syss]
            c.relax(opts)
            res[graph] = [(c.relax_status(i), c.instance_state(i)) for i in ids]
        finally:
            c.close()
    for (sa, xa), (sb, xb) in zip(res[False], res[True]):
        assert sa == sb and sa["state"] == "converged"
        for a, b in zip(xa, xb):
            assert np.array_equal(a, b)


def test_tables_loaded_between_ticks(tmp_path):
    """Tables loaded while instances are resident take effect at the next tick,
    and the captured tick graphs, which carry the table pointers and grid in
    their kernel arguments, are captured again: after a few ticks with the
    reference tables (10,000 rows), a switch to a 3,000-row synthetic set and a
    reset of the unknowns to the start, the relaxation equals bit for bit the
    one of a context that had the synthetic tables from the start."""
    import dmdgpu.tables as tabs
    say("[gpu] tables loaded between ticks")
    tabs.synthetic_mgh(str(tmp_path), seed=3, nr=3000, ne=3000)
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=21)
    opts = dmdgpu.relax_opts(atol=1e-8, max_newton=4, max_cg=80, eta=0.05)
    a = dmdgpu.Context(0)
    b = dmdgpu.Context(0, str(tmp_path))
    try:
        ia = a.add_instance(sd)
        a.relax_start(ia, opts)
        for _ in range(3):
            a.relax_tick()  # graphs captured with the reference tables
        for comb in range(3):
            assert a.L.dmdgpu_init_adp_MgHx(a.h, comb, str(tmp_path).encode()) == 0
        a.instance_set_unknowns(ia, q=sd["q"], sigma=sd["sigma"], x=sd["x"])
        a.relax(opts)  # restarts the resident instance from its (reset) state
        ib = b.add_instance(sd)
        b.relax(opts)
        sa, sb = a.relax_status(ia), b.relax_status(ib)
        say(f"    {sa}")
        assert sa == sb and sa["evals"] > 10, (sa, sb)
        for u, v in zip(a.instance_state(ia), b.instance_state(ib)):
            assert np.array_equal(u, v)
    finally:
        a.close()
        b.close()
