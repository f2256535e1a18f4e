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


def test_jvp_packed_host_and_device(gpu):
    """dmdgpu_jvp_packed with device buffers (a MatShell over device Vecs)
    gives bit for bit the host-buffer product, which is dmdgpu_jvp's; a zero
    direction gives zeros and a non-finite one is an argument error."""
    import ctypes as C
    sd = lat.hcp_mgh(4, 3, 3, x_h=(0.0, 0.3), seed=19)
    n = sd["n"]
    print(f"[gpu] jvp host and device buffers n={n}", flush=True)
    gpu.set_precision("fp64")
    gpu.set_state(sd)
    gpu.neighbors()
    rng = np.random.default_rng(7)
    X = np.concatenate([sd["q"].ravel(), sd["sigma"]])
    V = np.concatenate([rng.normal(size=3 * n), 0.1 * rng.normal(size=n)])
    host = gpu.jvp_packed(X, V)
    jq, js = gpu.jvp(sd["q"], sd["sigma"], V[:3 * n].reshape(n, 3), V[3 * n:])
    assert np.array_equal(host, np.concatenate([jq.ravel(), js]))
    hip = C.CDLL("libamdhip64.so.7")
    ptr = [C.c_void_p() for _ in range(3)]
    for p in ptr:
        assert hip.hipMalloc(C.byref(p), C.c_size_t(8 * 4 * n)) == 0
    try:
        for p, a in zip(ptr[:2], (X, V)):
            assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * 4 * n), 1) == 0
        gpu.jvp_packed(device=(ptr[0].value, ptr[1].value, ptr[2].value))
        dev = np.empty(4 * n)
        assert hip.hipMemcpy(dev.ctypes.data_as(C.c_void_p), ptr[2], C.c_size_t(8 * 4 * n), 2) == 0
        assert np.array_equal(dev, host)
    finally:
        for p in ptr:
            hip.hipFree(p)
    assert not np.abs(gpu.jvp_packed(X, np.zeros(4 * n))).any()
    bad = V.copy()
    bad[5] = np.nan
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.jvp_packed(X, bad)
    assert e.value.code == 1
