"""GPU parity of the per-(u, i) hooks against the reference-order oracle.

dmdgpu_site_DV0_Dq and dmdgpu_site_DV0_Dxi give dV0_i/d(q_u, sigma_u) and
dV0_i/dx_u one (u, i) pair at a time, as the reference's dmd_equations entries
evaluate_DV0_i_Dq_u_MgHx (src/ADP/MgHx-mf-V-bulk.cpp:1823-2098) and
evaluate_DV0_i_Dxi_u_adp_MgHx (:2102-2325) do.  Checked against the oracle's
restatement of those two functions in reference order (O(nb^2) triplet loops,
the reference's Stroud rule and masks, oracle.ref_DV0_Dq / ref_DV0_Dxi) within
1e-10 of the largest entry, for u = i, every neighbour u of i, and sites that
are not neighbours (zeros); then, for every pair rule, the hooks summed with
the residual's weights equal the residual the path computes.
"""
import sys

import numpy as np
import pytest

from conftest import RC

import dmdgpu.lattice as lat

pytestmark = pytest.mark.gpu
TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def queries(nnb, nbr, centres, n, rng):
    us, is_ = [], []
    for i in centres:
        row = list(nbr[i, :nnb[i]])
        others = [v for v in rng.choice(n, 6, replace=False) if v != i and v not in row][:2]
        for u in [i] + row + others:
            us.append(int(u))
            is_.append(int(i))
    return np.array(us, dtype=np.int32), np.array(is_, dtype=np.int32)


@pytest.mark.parametrize("case", ["cluster", "periodic"])
def test_hooks_match_reference_order_oracle(gpu, oracle, case):
    if case == "cluster":
        sd = lat.cluster(3, 2, 2, x_h=(0.0, 0.4), seed=3)   # some H inactive (x < 5e-4)
    else:
        sd = lat.hcp_mgh(4, 3, 3, x_h=(0.0, 0.3), seed=8)
    say(f"[gpu] per-(u, i) hooks vs reference order, {case} ({sd['n']} sites)")
    st, nnb, nbr = oracle.neighbors(sd, RC)
    s = oracle.system(sd, nnb, nbr)
    rho = oracle.ref_mf_rho(s)
    rng = np.random.default_rng(4)
    hs = np.flatnonzero(sd["species"] == lat.H)
    centres = list(rng.choice(np.flatnonzero(sd["species"] == lat.MG), 4, replace=False)) + \
        list(rng.choice(hs, 4, replace=False)) + [int(hs[np.argmin(sd["x"][hs])])]
    us, is_ = queries(nnb, nbr, centres, sd["n"], rng)
    gpu.set_state(sd)
    gpu.neighbors()
    dq = gpu.site_DV0_Dq(us, is_)
    dx = gpu.site_DV0_Dxi(us, is_)
    ref_q = np.array([oracle.ref_DV0_Dq(s, int(u), int(i), rho) for u, i in zip(us, is_)])
    ref_x = np.array([oracle.ref_DV0_Dxi(s, int(u), int(i), rho) for u, i in zip(us, is_)])
    eq = np.abs(dq[:, :3] - ref_q[:, :3]).max() / np.abs(ref_q[:, :3]).max()
    es = np.abs(dq[:, 3] - ref_q[:, 3]).max() / np.abs(ref_q[:, 3]).max()
    ex = np.abs(dx - ref_x).max() / np.abs(ref_x).max()
    say(f"    {len(us)} queries: q {eq:.2e}, sigma {es:.2e}, x {ex:.2e}")
    assert eq < TOL and es < TOL and ex < TOL
    far = np.array([u != i and u not in nbr[i, :nnb[i]] for u, i in zip(us, is_)])
    assert far.any() and not np.abs(dq[far]).any() and not np.abs(dx[far]).any()


@pytest.mark.parametrize("rule", [3, 5, 44, 6])
def test_hooks_sum_to_residual(gpu, oracle, rule):
    """R_q(u) = sum_i beta_i dV0_i/dq_u, R_sigma(u) = sum_i beta_i dV0_i/dsigma_u
    - 3/sigma_u; R_x of an H site = sum_i beta_i dV0_i/dx_u + its entropy terms
    (the Mg-centre form of the x hook agrees with the exact derivative when the
    Mg occupancies are 1, as here)."""
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.4), seed=6)
    say(f"[gpu] hooks sum to the residual, rule {rule}")
    assert np.all(sd["x"][sd["species"] == lat.MG] == 1.0)
    gpu.set_state(sd)
    gpu.neighbors()
    gpu.set_rule(rule)
    try:
        R = gpu.residual()
        nn, lst = gpu.get_neighbors()
        us, is_ = [], []
        for u in range(sd["n"]):
            for i in [u] + list(lst[u, :nn[u]]):
                us.append(u)
                is_.append(int(i))
        us, is_ = np.array(us, dtype=np.int32), np.array(is_, dtype=np.int32)
        dq = gpu.site_DV0_Dq(us, is_)
        dx = gpu.site_DV0_Dxi(us, is_)
    finally:
        gpu.set_rule(3)
    b = sd["beta"][is_]
    n = sd["n"]
    Rq = np.zeros((n, 3))
    Rs = np.zeros(n)
    Rx = np.zeros(n)
    np.add.at(Rq, us, b[:, None] * dq[:, :3])
    np.add.at(Rs, us, b * dq[:, 3])
    np.add.at(Rx, us, b * dx)
    act = sd["x"] >= 5e-4
    Rs[act] -= 3.0 / sd["sigma"][act]
    x = sd["x"]
    ent = -1.5 / x - np.where(x < 1.0, 1.0 / (1.0 - x), 0.0) - \
        np.where((sd["species"] == lat.H) & (x < 0.9999), sd["gamma"], 0.0)
    h = (sd["species"] == lat.H) & act
    Rx[h] += ent[h]
    assert np.abs(Rq - R["Rq"]).max() < TOL * np.abs(R["Rq"]).max()
    assert np.abs(Rs - R["Rs"]).max() < TOL * np.abs(R["Rs"]).max()
    assert np.abs(Rx[h] - R["Rx"][h]).max() < TOL * np.abs(R["Rx"][h]).max()
