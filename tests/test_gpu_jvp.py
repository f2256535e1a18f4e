"""GPU test of the matrix-free Jacobian-vector product (dmdgpu_jvp).

The product is checked against a central finite difference of the oracle's
residual at the same points, and against the symmetry that the Hessian of
the free-entropy total must have, w.(J v) = v.(J w).  Symmetry needs R to be
the exact gradient, which the reference tables' end clamp breaks (the clamp
returns the end slope with the end value), so that check runs on smooth
synthetic tables that vanish with zero slope at the cut-off.
"""
import numpy as np
import pytest

from conftest import RC

import dmdgpu
import dmdgpu.lattice as lat
import dmdgpu.tables as tabs

pytestmark = pytest.mark.gpu


def packed(r):
    return np.concatenate([np.ravel(r["Rq"]), np.ravel(r["Rs"])])


def test_jvp_matches_oracle_fd(gpu, oracle):
    sd = lat.hcp_mgh(4, 3, 3, x_h=(0.0, 0.3), seed=17)
    n = sd["n"]
    gpu.set_precision("fp64")
    gpu.set_state(sd)
    gpu.neighbors()
    rng = np.random.default_rng(3)
    vq, vs = rng.normal(size=(n, 3)), 0.1 * rng.normal(size=n)
    jq, js = gpu.jvp(sd["q"], sd["sigma"], vq, vs)
    print(f"[gpu] jvp n={n}", flush=True)
    # oracle central difference at the same displacement
    st, nnb, nbr = oracle.neighbors(sd, RC)
    v = np.concatenate([vq.ravel(), vs])
    eps = 1e-5 * np.sqrt(4 * n / (v @ v))
    res = []
    for s in (eps, -eps):
        t = dict(sd)
        t["q"] = sd["q"] + s * vq
        t["sigma"] = sd["sigma"] + s * vs
        res.append(packed(oracle.fact_residual(oracle.system(t, nnb, nbr))))
    ref = (res[0] - res[1]) / (2 * eps)
    got = np.concatenate([jq.ravel(), js])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"    max rel err vs oracle FD {err:.2e}", flush=True)
    assert err < 1e-6
    # the state is left at X
    r0 = gpu.residual()
    t = oracle.fact_residual(oracle.system(sd, nnb, nbr))
    assert np.abs(r0["Rq"] - t["Rq"]).max() <= 1e-10 * np.abs(t["Rq"]).max()


def test_jvp_hessian_symmetry_on_smooth_tables(tmp_path):
    tabs.synthetic_mgh(str(tmp_path), seed=3, nr=3000, ne=3000)
    ctx = dmdgpu.Context(0, str(tmp_path), 6.5)
    try:
        sd = lat.hcp_mgh(5, 3, 3, x_h=(0.05, 0.5), seed=21)
        n = sd["n"]
        ctx.set_state(sd)
        ctx.neighbors()
        rng = np.random.default_rng(5)
        vq, vs = rng.normal(size=(n, 3)), 0.1 * rng.normal(size=n)
        wq, ws = rng.normal(size=(n, 3)), 0.1 * rng.normal(size=n)
        jq, js = ctx.jvp(sd["q"], sd["sigma"], vq, vs)
        kq, ks = ctx.jvp(sd["q"], sd["sigma"], wq, ws)
        a = np.concatenate([wq.ravel(), ws]) @ np.concatenate([jq.ravel(), js])
        b = np.concatenate([vq.ravel(), vs]) @ np.concatenate([kq.ravel(), ks])
        print(f"[gpu] smooth tables: w.Jv {a:.12e}  v.Jw {b:.12e}", flush=True)
        assert abs(a - b) <= 1e-7 * max(abs(a), abs(b))
    finally:
        ctx.close()
