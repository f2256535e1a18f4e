"""CPU tests that pin the oracle's diagonal Hessian blocks (no GPU).

evaluate_D2V0_i_Dmeanq2_u_MgHx (src/ADP/MgHx-mf-V-bulk.cpp:1366-1599) has no
test or golden vector in the reference, so its restatement is pinned by:
  * the analytic Hessians of the four ADP functions against second differences
    of their values (the values are pinned by the energy tests);
  * the centre blocks d2V0_u/dq_u2 against finite differences of the reference-
    order gradient dV0_u/dq_u (the density is re-evaluated at each step, so the
    F'' g g^T term is covered);
  * the neighbour blocks d2V0_i/dq_u2 against the same finite differences less
    the mixed codes 5 + 7 of the coincident triplet (i, u, u), which that
    function does not add (it sums codes 4 and 8 of u, :1540-1570); the mixed
    codes come from the function Hessians pinned above, summed over the
    enumerated 12 points of the 6D rule;
  * the factorised form (the one the GPU runs) against the reference order;
  * the factorised exact blocks against finite differences of the factorised
    force residual, for the 3rd-order and the 5th-order pair rules.
"""
import numpy as np
import pytest

from conftest import RC

import dmdgpu.lattice as lat
import dmdgpu.tables as tabs
from oracle.oracle import Oracle

SQRT6 = np.sqrt(6.0)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def synth(tmp_path_factory):
    d = tmp_path_factory.mktemp("synth_h")
    tabs.synthetic_mgh(str(d), seed=3, nr=3000, ne=3000)
    return Oracle(str(d))


def _copy(sysd):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sysd.items()}


@pytest.mark.parametrize("which,ns", [(0, 2), (1, 2), (2, 3), (3, 3)])
def test_function_hessians_match_second_differences(oracle, which, ns):
    """d2F/dq_a dq_b of rho_ij, V_pair_ij, V_dipole_ij1j2, V_quadrupole_ij1j2
    (every code) against central second differences of F, at distances inside
    the tables (2.2 - 5.5 A)."""
    rng = np.random.default_rng(7 + which)
    h = 1e-4
    for trial in range(3):
        q = np.zeros((ns, 3))
        q[0] = rng.normal(size=3)
        for k in range(1, ns):
            d = rng.normal(size=3)
            q[k] = q[0] + d / np.linalg.norm(d) * rng.uniform(2.2, 5.5)
        n = rng.uniform(0.3, 1.0, ns)
        slots = rng.integers(0, 2, ns)
        for code in range(ns * ns):
            a, b = divmod(code, ns)
            H = oracle.fn_eval(which, 2, code, n, q, slots)
            fd = np.zeros((3, 3))
            for x in range(3):
                for y in range(3):
                    def f(dx, dy):
                        qq = q.copy()
                        qq[a, x] += dx
                        qq[b, y] += dy
                        return oracle.fn_eval(which, 0, 0, n, qq, slots)
                    fd[x, y] = (f(h, h) - f(h, -h) - f(-h, h) + f(-h, -h)) / (4 * h * h)
            assert np.abs(H - fd).max() < 1e-5 * np.abs(H).max() + 1e-12, (which, code)


def _grad_fd(o, sysd, nnb, nbr, u, i, h=1e-5):
    """d/dq_u of the reference-order dV0_i/dq_u, density re-evaluated."""
    fd = np.zeros((3, 3))
    for c in range(3):
        out = []
        for sg in (1, -1):
            sd = _copy(sysd)
            sd["q"][u, c] += sg * h
            s = o.system(sd, nnb, nbr)
            out.append(o.ref_DV0_Dq(s, u, i, o.ref_mf_rho(s))[:3])
        fd[:, c] = (out[0] - out[1]) / (2 * h)
    return fd


def _coincident_mixed(o, sysd, u, i):
    """Codes 5 + 7 of dipole + quadrupole on the triplet (i, u, u) under the
    6D rule (:1516-1520 couples j1 = j2): points q_i +- sqrt(6) sigma_i e_a or
    both copies of q_u +- sqrt(6) sigma_u e_a, weight 1/12 (free cluster, so
    no periodic images)."""
    q0 = np.stack([sysd["q"][i], sysd["q"][u], sysd["q"][u]])
    n = np.array([sysd["x"][i], sysd["x"][u], sysd["x"][u]])
    sl = np.array([int(sysd["species"][k] == lat.H) for k in (i, u, u)])
    acc = np.zeros((3, 3))
    for row, sig, sites in ((0, sysd["sigma"][i], (0,)), (1, sysd["sigma"][u], (1, 2))):
        for a in range(3):
            for sg in (1.0, -1.0):
                q = q0.copy()
                for st in sites:
                    q[st, a] += sg * SQRT6 * sig
                for which in (2, 3):
                    acc += (o.fn_eval(which, 2, 5, n, q, sl) + o.fn_eval(which, 2, 7, n, q, sl)) / 12
    return acc


def test_reference_blocks_against_gradient_differences(synth):
    sysd = lat.cluster(2, 1, 1, x_h=(0.05, 0.5), seed=4, sigma_jitter=0.3)
    st, nnb, nbr = synth.neighbors(sysd, 6.5)
    assert st == 0
    s = synth.system(sysd, nnb, nbr)
    mf = synth.ref_mf_rho(s)
    rng = np.random.default_rng(3)
    visible = 0
    for u in rng.choice(sysd["n"], 3, replace=False):
        H = synth.ref_D2V0_Dq2(s, u, u, mf)
        fd = _grad_fd(synth, sysd, nnb, nbr, u, u)
        assert rel(H, fd) < 1e-8
        assert np.abs(H - H.T).max() < 1e-12 * np.abs(H).max()
        for i in nbr[u, :3]:
            H = synth.ref_D2V0_Dq2(s, u, i, mf)
            fd = _grad_fd(synth, sysd, nnb, nbr, u, i)
            mixed = _coincident_mixed(synth, sysd, u, i)
            visible += np.abs(mixed).max() > 1e-3 * np.abs(H).max()
            assert rel(H + mixed, fd) < 1e-7
    assert visible >= 2  # the omitted codes matter on some pairs


@pytest.mark.parametrize("maker", [lambda: lat.cluster(2, 1, 1, x_h=(0.0, 0.4), seed=3),
                                   lambda: lat.hcp_mgh(4, 3, 3, x_h=(0.0, 0.3), seed=9)])
def test_factorised_blocks_equal_reference_order(oracle, maker):
    sysd = maker()
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    s = oracle.system(sysd, nnb, nbr)
    mf = oracle.ref_mf_rho(s)
    F = oracle.fact_diag_blocks(s)
    n = sysd["n"]
    sites = np.arange(n) if n < 100 else np.random.default_rng(0).choice(n, 10, replace=False)
    R = oracle.ref_diag_blocks(s, mf, sites)
    assert rel(F[sites], R) < 1e-12


@pytest.mark.parametrize("rule", [3, 5, 44, 6])
def test_exact_blocks_are_residual_derivatives(synth, rule):
    sysd = lat.hcp_mgh(4, 3, 3, x_h=(0.05, 0.5), seed=21)
    st, nnb, nbr = synth.neighbors(sysd, 6.5)
    D = synth.fact_diag_blocks(synth.system(sysd, nnb, nbr, rule), exact=True)
    rng = np.random.default_rng(5)
    h = 1e-5
    for u in rng.choice(sysd["n"], 3, replace=False):
        fd = np.zeros((3, 3))
        for c in range(3):
            out = []
            for sg in (1, -1):
                sd = _copy(sysd)
                sd["q"][u, c] += sg * h
                out.append(synth.fact_residual(synth.system(sd, nnb, nbr, rule))["Rq"][u])
            fd[:, c] = (out[0] - out[1]) / (2 * h)
        assert rel(D[u], fd) < 1e-7


def test_masked_sites_have_zero_blocks(oracle):
    sysd = lat.cluster(2, 1, 1, x_h=(0.0, 0.3), seed=11)
    hs = np.nonzero(sysd["species"] == lat.H)[0]
    sysd["x"][hs[:3]] = [0.0, 1e-4, 4.99e-4]
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    s = oracle.system(sysd, nnb, nbr)
    F = oracle.fact_diag_blocks(s)
    assert (F[hs[:3]] == 0).all()
    R = oracle.ref_diag_blocks(s, oracle.ref_mf_rho(s))
    assert rel(F, R) < 1e-12


# ---------------------------------------------------------------------------
# Zero temperature (evaluate_DV_i_Dq_u_adp_MgHx :147-329 and
# evaluate_D2V_i_Dq2_u_MgHx :333-526): the functions of the mean-field forms
# evaluated at the mean positions, so the sigma -> 0 limits of the pinned
# mean-field residual and blocks; the Hessian also against differences of the
# zero-temperature gradient.
# ---------------------------------------------------------------------------

def test_zero_temperature_force_and_blocks_are_width_limits(oracle):
    sysd = lat.cluster(2, 1, 1, x_h=(0.0, 0.4), seed=3)
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    errs = []
    for sg in (1e-5, 1e-6):
        sysd["sigma"][:] = sg
        s = oracle.system(sysd, nnb, nbr)
        F, D = oracle.ref_zero_t(s, oracle.ref_rho0(s))
        mf = oracle.ref_mf_rho(s)
        Rq, _, _ = oracle.ref_residual(s, mf)
        errs.append((rel(Rq, F), rel(oracle.ref_diag_blocks(s, mf), D)))
    assert errs[1][0] < 2e-2 * errs[0][0] + 1e-14  # O(sigma^2)
    assert errs[1][0] < 1e-9 and errs[1][1] < 1e-9


def test_zero_temperature_hessian_against_gradient_differences(synth):
    sysd = lat.cluster(2, 1, 1, x_h=(0.05, 0.5), seed=4, sigma_jitter=0.3)
    st, nnb, nbr = synth.neighbors(sysd, 6.5)
    s = synth.system(sysd, nnb, nbr)
    rho0 = synth.ref_rho0(s)
    rng = np.random.default_rng(8)
    h = 1e-5
    for u in rng.choice(sysd["n"], 3, replace=False):
        for i in [u] + list(nbr[u, :2]):
            H = synth.ref_D2V_Dq2_zero_t(s, u, i, rho0)
            fd = np.zeros((3, 3))
            for c in range(3):
                out = []
                for sg in (1, -1):
                    sd = _copy(sysd)
                    sd["q"][u, c] += sg * h
                    ss = synth.system(sd, nnb, nbr)
                    out.append(synth.ref_DV_Dq_zero_t(ss, u, i, synth.ref_rho0(ss)))
                fd[:, c] = (out[0] - out[1]) / (2 * h)
            if i != u:
                # codes 5 + 7 of the coincident triplet (i, u, u) at the mean point
                q = np.stack([sysd["q"][i], sysd["q"][u], sysd["q"][u]])
                n = np.array([sysd["x"][i], sysd["x"][u], sysd["x"][u]])
                sl = np.array([int(sysd["species"][k] == lat.H) for k in (i, u, u)])
                for which in (2, 3):
                    H = H + synth.fn_eval(which, 2, 5, n, q, sl) + synth.fn_eval(which, 2, 7, n, q, sl)
            assert rel(H, fd) < 1e-7
