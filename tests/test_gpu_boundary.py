"""GPU tests of the C-ABI boundary's guards.

* A neighbour build that fails leaves the instance without lists: no
  evaluation reads truncated rows (DMDGPU_ERR_STATE until a build succeeds).
* The device-memory budget bounds every allocation of the context, also while
  the arena grows: the context's own count and the device's free memory
  (hipMemGetInfo) both stay within it across admissions.
* A density at or below -drho of the embedding table (the reference's spline
  indexes a negative segment there, src/Numerical/cubic-spline.cpp:84-93) is
  reported as DMDGPU_ERR_DOMAIN exactly when the oracle's densities put a
  site there; a relaxation that meets it ends in the "domain" state.
* Non-finite positions, widths or occupancies are argument errors, and so is
  an instance above the 6,710,784-site limit.
* An arena growth that fails part way (forced from the host) is an error of
  the add that asked for it and leaves the resident instances intact.
* Host writes (state updates, added instances) wait for an evaluation still
  in flight on the context's stream.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import RC

import dmdgpu
import dmdgpu.lattice as lat
import dmdgpu.tables as tab

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_HIP, ERR_MAXNEIGH, ERR_STATE, ERR_BUDGET, ERR_DOMAIN = 1, 2, 5, 8, 10, 11


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def hip_free_bytes():
    """Free device memory from the HIP runtime the library itself uses."""
    dmdgpu.load_library()
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_failed_neighbor_build_invalidates_lists(gpu):
    say("[gpu] failed neighbour build -> ERR_STATE")
    sd = lat.cluster(6, 4, 4, x_h=(0.05, 0.3), seed=3)  # 1152 sites, free
    gpu.set_state(sd)
    gpu.neighbors()
    gpu.residual()  # a good build evaluates
    ctr = sd["q"].mean(0)
    dense = ctr + 0.45 * (sd["q"] - ctr)  # ~11x the density: > 250 neighbours
    gpu.update_state(q=dense)
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.neighbors()
    assert e.value.code == ERR_MAXNEIGH
    for call in (gpu.residual, gpu.mf_rho, gpu.site_energies):
        with pytest.raises(dmdgpu.DmdError) as e:
            call()
        assert e.value.code == ERR_STATE
    gpu.update_state(q=sd["q"])
    gpu.neighbors()  # lists valid again
    gpu.residual()


def test_budget_bounds_device_memory():
    say("[gpu] budget bounds every device allocation")
    budget = 512 * 2 ** 20
    ctx = dmdgpu.Context(0)
    try:
        free0 = hip_free_bytes()
        ctx.set_budget(budget)
        ids, refused = [], 0
        for k in range(40):
            sd = lat.box_of_size(6000 + 500 * k, seed=k)
            try:
                ids.append(ctx.add_instance(sd))
            except dmdgpu.DmdError as e:
                assert e.code == ERR_BUDGET
                refused += 1
                if refused >= 3:
                    break
            held = ctx.device_bytes()
            used = free0 - hip_free_bytes()
            say(f"    {len(ids)} instances, {held / 2**20:.1f} MiB held, "
                f"{used / 2**20:.1f} MiB taken from the device")
            assert held <= budget
            # allocator rounding: at most 2 MiB per allocation (about 60 arrays)
            assert used <= budget + 128 * 2 ** 20
        assert refused > 0 and len(ids) >= 2
        ctx.evaluate()  # what was admitted still evaluates
        for i in ids[:2]:
            ctx.instance_residual(i)
    finally:
        ctx.close()


@pytest.mark.parametrize("k", [1, 14, 39, 47])
def test_failed_arena_growth_keeps_resident_instances(monkeypatch, k):
    """DMDGPU_TEST_FAIL_GROWTH = k refuses the k-th per-site array's allocation
    in the next arena growth, as an out-of-memory device would: k = 1 the
    positions (contents kept, nothing grown yet), 14 the row slots, 39 the
    diagonal blocks and 47 the last array (no contents kept: released before
    their allocation, so they must be put back).  The add fails with
    DMDGPU_ERR_HIP; the resident instance evaluates bit for bit as before; the
    same add then succeeds and equals its evaluation in a context of its own."""
    say(f"[gpu] failed arena growth at array {k}")
    small, big = lat.box_of_size(3000, seed=1), lat.box_of_size(30000, seed=2)
    ctx, alone = dmdgpu.Context(0), dmdgpu.Context(0)
    try:
        ia = ctx.add_instance(small)
        ctx.evaluate()
        ra = ctx.instance_residual(ia)
        monkeypatch.setenv("DMDGPU_TEST_FAIL_GROWTH", str(k))
        with pytest.raises(dmdgpu.DmdError) as e:
            ctx.add_instance(big)
        assert e.value.code == ERR_HIP
        monkeypatch.delenv("DMDGPU_TEST_FAIL_GROWTH")
        ctx.evaluate()
        r = ctx.instance_residual(ia)
        for key in ("Rq", "Rs", "Rx"):
            assert np.array_equal(r[key], ra[key]), key
        ib = ctx.add_instance(big)
        ctx.evaluate()
        rb = ctx.instance_residual(ib)
        r = ctx.instance_residual(ia)
        alone.add_instance(big)
        alone.evaluate()
        rb1 = alone.instance_residual(0)
        for key in ("Rq", "Rs", "Rx"):
            assert np.array_equal(r[key], ra[key]), key
            assert np.array_equal(rb[key], rb1[key]), key
    finally:
        ctx.close()
        alone.close()


def test_table_domain_status(oracle, tmp_path):
    """Tables whose Mg density is negative drive rho_bar below -drho of the Mg
    embedding table: every call that reads an evaluation reports it, as the
    oracle's densities say it should; the same system with the positive
    synthetic tables evaluates normally."""
    from oracle.oracle import Oracle
    say("[gpu] table-domain status")
    sd = lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=5)
    for sign in (1.0, -1.0):
        folder = str(tmp_path / f"tables{int(sign)}")
        tab.synthetic_mgh(folder, seed=2)
        if sign < 0:  # rewrite adpmg.dat with the Mg density negated
            rng = np.random.default_rng(2)
            jit = lambda: 1.0 + 0.1 * (2 * rng.random() - 1)
            fn = {"rho_Mg": tab.smooth(-0.05 * jit(), 1.0 * jit(), 3.2, 6.5)}
            path = folder + "/adpmg.dat"
            lines = open(path).read().splitlines()
            nr = int(lines[1].split()[0])
            dx = float(lines[1].split()[1])
            rows = tab.hermite_rows(*fn["rho_Mg"], nr, dx)
            lines[2:2 + nr] = [" ".join(f"{v:.17e}" for v in r) for r in rows]
            open(path, "w").write("\n".join(lines) + "\n")
        O = Oracle(folder)
        st, nnb, nbr = O.neighbors(sd, 6.5)
        ref = O.fact_residual(O.system(sd, nnb, nbr))
        # embedding-table step of each site's species (slot 0 Mg, 1 H)
        dxe = np.where(sd["species"] == lat.H, O.spline_size(1, 1)[1], O.spline_size(1, 0)[1])
        act = sd["x"] >= 5e-4
        expect = bool((act & (ref["mf_rho"] <= -dxe)).any())
        assert expect == (sign < 0)
        ctx = dmdgpu.Context(0, folder, 6.5)
        try:
            ctx.set_state(sd)
            ctx.neighbors()
            if not expect:
                ctx.residual()
                continue
            for call in (ctx.residual, ctx.mf_rho, ctx.site_energies, ctx.diag_blocks):
                with pytest.raises(dmdgpu.DmdError) as e:
                    call()
                assert e.value.code == ERR_DOMAIN
            ctx.relax(dmdgpu.relax_opts(atol=1e-8, max_newton=5))
            assert ctx.relax_status(0)["state"] == "domain"
        finally:
            ctx.close()


def test_nonfinite_inputs_rejected(gpu):
    say("[gpu] non-finite inputs")
    sd = lat.cluster(2, 2, 1, x_h=(0.05, 0.3), seed=1)
    for key in ("q", "sigma", "x"):
        bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sd.items()}
        bad[key].flat[3] = np.nan
        with pytest.raises(dmdgpu.DmdError) as e:
            gpu.set_state(bad)
        assert e.value.code == ERR_ARG, key
    gpu.set_state(sd)
    gpu.neighbors()
    q = sd["q"].copy()
    q[0, 0] = np.inf
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.update_state(q=q)
    assert e.value.code == ERR_ARG


def test_instance_size_limit(gpu):
    """An instance above 6,710,784 sites (its adjoint records would pass the
    4 GiB that the gradient kernel's 32-bit row offsets reach) is an argument
    error, refused before anything is read or allocated."""
    say("[gpu] instance size limit")
    n = 6_710_784 + 1
    big = {"n": n, "q": np.zeros((n, 3)), "sigma": np.full(n, 0.1), "x": np.ones(n),
           "species": np.full(n, 12, dtype=np.int32), "beta": np.ones(n), "gamma": np.zeros(n),
           "lo": np.zeros(3), "hi": np.full(3, 1e4), "periodic": np.zeros(3, dtype=np.int32)}
    ctx = dmdgpu.Context(0)
    try:
        before = ctx.device_bytes()
        with pytest.raises(dmdgpu.DmdError) as e:
            ctx.add_instance(big)
        assert e.value.code == ERR_ARG
        assert ctx.device_bytes() == before
    finally:
        ctx.close()


def test_residual_packed_host_and_device(gpu, oracle):
    """The packed SNES FormFunction form: host buffers and device buffers give
    bit for bit the residual of dmdgpu_residual at X, in the caller's order,
    within 1e-10 of the oracle; a width that is not positive is an argument
    error."""
    say("[gpu] packed residual, host and device buffers")
    sd = lat.hcp_mgh(5, 3, 3, x_h=(0.0, 0.3), seed=9)
    n = sd["n"]
    gpu.set_state(sd)
    gpu.neighbors()
    rng = np.random.default_rng(3)
    q = sd["q"] + rng.normal(0, 0.02, sd["q"].shape)
    sg = sd["sigma"] * (1 + 0.05 * rng.random(n))
    X = np.concatenate([q.ravel(), sg])
    R, tot = gpu.residual_packed(X)
    gpu.update_state(q=q, sigma=sg)
    ref = gpu.residual()
    assert np.array_equal(R[:3 * n].reshape(n, 3), ref["Rq"])
    assert np.array_equal(R[3 * n:4 * n], ref["Rs"])
    assert np.array_equal(R[4 * n:], ref["Rx"])
    assert tot[0] == ref["S0_total"] and tot[1] == ref["V0_total"]
    fin = dict(sd)
    fin["q"], fin["sigma"] = q, sg
    nnb, nbr = gpu.get_neighbors()  # the lists of the start state, kept by the update
    o = oracle.fact_residual(oracle.system(fin, nnb, nbr))
    assert np.abs(R[:3 * n] - o["Rq"].ravel()).max() < 1e-10 * np.abs(o["Rq"]).max()
    # device buffers (same HIP runtime as the library)
    hip = C.CDLL("libamdhip64.so.7")
    dX, dR = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dX), C.c_size_t(8 * 4 * n)) == 0
    assert hip.hipMalloc(C.byref(dR), C.c_size_t(8 * 5 * n)) == 0
    try:
        X2 = np.concatenate([sd["q"].ravel(), sd["sigma"]])
        assert hip.hipMemcpy(dX, X2.ctypes.data_as(C.c_void_p), C.c_size_t(8 * 4 * n), 1) == 0
        _, tot2 = gpu.residual_packed(device=(dX.value, dR.value))
        Rd = np.empty(5 * n)
        assert hip.hipMemcpy(Rd.ctypes.data_as(C.c_void_p), dR, C.c_size_t(8 * 5 * n), 2) == 0
        Rh, toth = gpu.residual_packed(X2)
        assert np.array_equal(Rd, Rh) and np.array_equal(tot2, toth)
    finally:
        hip.hipFree(dX)
        hip.hipFree(dR)
    bad = X.copy()
    bad[3 * n + 5] = -0.1
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.residual_packed(bad)
    assert e.value.code == ERR_ARG


def test_host_writes_wait_for_evaluations_in_flight():
    """dmdgpu_residual_resident returns with its kernels still running; a state
    update and an added instance issued right after it must wait for them (the
    library copies through the context's stream).  Each result equals the one
    of a context that synchronised in between."""
    say("[gpu] host writes after an evaluation in flight")
    a = lat.box_of_size(20000, seed=3)
    q2 = a["q"] + 0.01 * np.random.default_rng(4).standard_normal(a["q"].shape)
    b = lat.box_of_size(6000, seed=5)
    results = []
    for wait in (False, True):
        ctx = dmdgpu.Context(0)
        try:
            ctx.set_state(a)
            ctx.neighbors()
            for _ in range(3):
                ctx.residual_resident()
            if wait:
                ctx.sync()
            ctx.update_state(q=q2)  # overwrites positions the evaluation reads
            r = ctx.residual()
            ctx.residual_resident()
            if wait:
                ctx.sync()
            ib = ctx.add_instance(b)  # grows the arena the evaluation uses
            ctx.evaluate()
            results.append((r, ctx.instance_residual(ib)))
        finally:
            ctx.close()
    for k in ("Rq", "Rs", "Rx"):
        assert np.array_equal(results[0][0][k], results[1][0][k]), k
        assert np.array_equal(results[0][1][k], results[1][1][k]), k
