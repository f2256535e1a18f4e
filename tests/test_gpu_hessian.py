"""GPU parity of the diagonal Hessian blocks (dmdgpu_diag_blocks) through the
C ABI.

Bars (FP64): within 1e-10 relative (max-norm of the difference over max-norm
of the oracle) of the reference-order restatement of
evaluate_D2V0_i_Dmeanq2_u_MgHx (MgHx-mf-V-bulk.cpp:1366-1599) summed over
i in {u} U N(u), and of the factorised restatement for both assemblies and
both pair rules.  At the 54,000-site size, where the oracle is slow, the exact
blocks are checked against central differences of the GPU's own force
residual (whose parity with the oracle the residual tests hold).
"""
import sys

import numpy as np
import pytest

from conftest import RC

import dmdgpu.lattice as lat

pytestmark = pytest.mark.gpu
TOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(gpu, oracle, sysd, rule=3):
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    assert st == 0
    gpu.set_state(sysd)
    gpu.set_tuning(8, 16, 0)
    gpu.set_precision("fp64")
    gpu.set_rule(rule)
    gpu.neighbors()
    return oracle.system(sysd, nnb, nbr, rule)


@pytest.mark.parametrize("maker", [lambda: lat.cluster(2, 1, 1, x_h=(0.0, 0.4), seed=3),
                                   lambda: lat.hcp_mgh(4, 3, 3, x_h=(0.0, 0.3), seed=9)])
def test_diag_blocks_match_reference_order(gpu, oracle, maker):
    sysd = maker()
    say(f"[gpu] diag blocks vs reference order, {sysd['n']} sites")
    s = _setup(gpu, oracle, sysd)
    D = gpu.diag_blocks()
    n = sysd["n"]
    sites = np.arange(n) if n < 100 else np.random.default_rng(0).choice(n, 12, replace=False)
    R = oracle.ref_diag_blocks(s, oracle.ref_mf_rho(s), sites)
    err = rel(D[sites], R)
    say(f"    max rel err {err:.3e}")
    assert err < TOL


@pytest.mark.parametrize("rule", [3, 5, 44, 6])
@pytest.mark.parametrize("exact", [False, True])
def test_diag_blocks_match_factorised(gpu, oracle, rule, exact):
    sysd = lat.hcp_mgh(5, 3, 3, x_h=(0.0, 0.3), seed=9)
    say(f"[gpu] diag blocks vs factorised, rule {rule}, exact {exact}")
    try:
        s = _setup(gpu, oracle, sysd, rule)
        D = gpu.diag_blocks(exact=exact)
    finally:
        gpu.set_rule(3)
    F = oracle.fact_diag_blocks(s, exact=exact)
    err = rel(D, F)
    say(f"    max rel err {err:.3e}")
    assert err < TOL
    assert np.abs(D - D.transpose(0, 2, 1)).max() <= 1e-13 * np.abs(D).max()


def test_diag_blocks_are_fp64_under_split_precision_and_zero_when_masked(gpu, oracle):
    sysd = lat.cluster(3, 2, 2, x_h=(0.0, 0.4), seed=5)
    hs = np.nonzero(sysd["species"] == lat.H)[0]
    sysd["x"][hs[:3]] = [0.0, 1e-4, 4.99e-4]
    say("[gpu] diag blocks under split precision, masked sites")
    s = _setup(gpu, oracle, sysd)
    D64 = gpu.diag_blocks()
    try:
        gpu.set_precision("split32")
        D32 = gpu.diag_blocks()
        gpu.residual()  # the split evaluation still runs after the FP64 blocks
    finally:
        gpu.set_precision("fp64")
    assert np.array_equal(D64, D32)
    assert (D64[hs[:3]] == 0).all()
    assert rel(D64, oracle.fact_diag_blocks(s)) < TOL


def test_exact_blocks_are_residual_derivatives_at_full_size(tmp_path):
    """54,000 sites (20x15x15 periodic, the headline workload's size): column a
    of the exact D_u equals the central difference of the GPU's own R_q(u) in
    q_(u,a) (h = 1e-5 A).  Smooth synthetic tables: the reference tables are
    cubic splines whose first derivative jumps by up to 2e-6 at knots (pair
    table), so a difference across a knot is off by ~1e-5 relative there; the
    smooth tables leave truncation and rounding, ~1e-9.  The reference
    assembly differs from the difference by the omitted coincident mixed codes."""
    import dmdgpu
    import dmdgpu.tables as tabs
    tabs.synthetic_mgh(str(tmp_path), seed=3, nr=3000, ne=3000)
    sysd = lat.hcp_mgh(20, 15, 15, x_h=(0.0, 0.1), seed=2024)
    say(f"[gpu] exact diag blocks vs residual differences, {sysd['n']} sites, smooth tables")
    g = dmdgpu.Context(0, str(tmp_path), 6.5)
    try:
        g.set_state(sysd)
        g.neighbors()
        D = g.diag_blocks(exact=True)
        Dr = g.diag_blocks(exact=False)
        assert np.isfinite(D).all()
        assert np.abs(D - D.transpose(0, 2, 1)).max() <= 1e-13 * np.abs(D).max()
        rng = np.random.default_rng(1)
        h = 1e-5
        sites = np.concatenate([rng.choice(np.nonzero(sysd["species"] == s_)[0], 2, replace=False)
                                for s_ in (lat.MG, lat.H)])
        for u in sites:
            if sysd["x"][u] < 5e-4:
                continue
            fd = np.zeros((3, 3))
            for a in range(3):
                col = []
                for sg in (1.0, -1.0):
                    q = sysd["q"].copy()
                    q[u, a] += sg * h
                    g.update_state(q=q)
                    col.append(g.residual()["Rq"][u])
                fd[:, a] = (col[0] - col[1]) / (2 * h)
            g.update_state(q=sysd["q"])
            err, err_r = rel(D[u], fd), rel(Dr[u], fd)
            say(f"    site {u} (species {sysd['species'][u]}): exact {err:.2e}, "
                f"reference assembly {err_r:.2e}")
            assert err < 1e-7
            assert err_r > 1e-5
    finally:
        g.close()


@pytest.mark.parametrize("maker", [lambda: lat.cluster(3, 2, 2, x_h=(0.0, 0.4), seed=3),
                                   lambda: lat.hcp_mgh(5, 3, 3, x_h=(0.0, 0.3), seed=9)])
def test_zero_temperature_force_and_blocks(gpu, oracle, maker):
    """sigma -> 0: the GPU force residual R_q and the diagonal blocks equal the
    reference's zero-temperature gradient and Hessian (MgHx-mf-V-bulk.cpp:
    147-329, 333-526) summed with the residual's weights.  sigma = 1e-12 A:
    the quadrature points sit 2.4e-12 A from the mean, so no point pair
    straddles a spline knot (where the reference tables' derivatives jump)
    and the O(sigma^2) rule error is below rounding."""
    sysd = maker()
    sysd["sigma"][:] = 1e-12
    say(f"[gpu] zero-temperature force and blocks, {sysd['n']} sites")
    s = _setup(gpu, oracle, sysd)
    F, D = oracle.ref_zero_t(s, oracle.ref_rho0(s))
    Rq = gpu.residual()["Rq"]
    G = gpu.diag_blocks()
    ef, ed = rel(Rq, F), rel(G, D)
    say(f"    force {ef:.2e}, blocks {ed:.2e}")
    assert ef < TOL and ed < TOL


def test_instance_blocks_in_a_batch_equal_alone(gpu, oracle):
    """dmdgpu_instance_diag_blocks on resident instances of a batch: each equals
    the blocks of the same system alone (bit for bit) and the oracle."""
    syss = [lat.cluster(2, 1, 1, x_h=(0.0, 0.4), seed=s) for s in (3, 4)]
    say("[gpu] instance diag blocks in a batch")
    alone = []
    for sd in syss:
        _setup(gpu, oracle, sd)
        alone.append(gpu.diag_blocks(exact=True))
    gpu.remove_instance(0)
    ids = [gpu.add_instance(sd) for sd in syss]
    for k, iid in enumerate(ids):
        D = gpu.instance_diag_blocks(iid, exact=True)
        assert np.array_equal(D, alone[k])
        st, nnb, nbr = oracle.neighbors(syss[k], RC)
        F = oracle.fact_diag_blocks(oracle.system(syss[k], nnb, nbr), exact=True)
        assert rel(D, F) < TOL
    for iid in ids:
        gpu.remove_instance(iid)
