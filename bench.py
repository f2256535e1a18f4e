"""Benchmark of the mean-field DMD force path on one MI355X.

A step is one full FP64 residual evaluation (phase-1 site sums, phase-2
gradient gather, deterministic totals) of the 54,000-site Mg-H workload:
hcp Mg 20 x 15 x 15 orthorhombic cells (18,000 Mg) plus 36,000 tetrahedral H
sites with occupancy x ~ U[0, 0.1], fully periodic, reference Mg-H ADP tables,
3rd-order Stroud quadrature.  Inputs are resident in HBM before timing starts.

Prints one JSON line: throughput in pair-quadrature evaluations per second
(PQE/s; 50 PQE per active ordered neighbour pair per residual), the roofline
figure of the dominant kernel (FP64 VALU) and the CPU baseline (the factorised
FP64 oracle on the host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bt-am_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# FP64 floating-point operations per active ordered pair of each phase of the
# FP64 evaluation (FMA = 2), counted once with the hardware counter
# SQ_INSTS_VALU_FLOPS_FP64 (x 64 lanes) on this workload and frozen here
# (profiles/valu_mgh_20x15x15.md).  Phase 1: site sums plus the pair-rule pass
# (once per unordered pair); phase 2: the pair kernel (each unordered pair
# once) plus the slot sums.  The PQE count per residual stays the workload's
# fixed 50 per active ordered pair.  (The pair-rule terms in radial form took
# these from 1771 and 2941 with the same results.)
FLOP_PER_PAIR = {"phase1_site_sums": 1480, "phase2_gradient": 2426}
# The split-precision evaluation, counted the same way (SQ_INSTS_VALU_FLOPS_FP32
# and _FP64 x 64 per active ordered pair, profiles/valu_mgh_20x15x15.md):
# (FP32, FP64) FLOPs per pair of each phase.
FLOP_PER_PAIR_SPLIT = {"phase1_site_sums": (712 + 351, 335 + 78),
                       "phase2_gradient": (2135, 229 + 22)}
# Relaxation ticks (Mg positions and widths, residual-norm line search) form
# less: the pair-rule pass the densities only, phase 1 no energy sums, phase 2
# no occupancy terms and only the pairs with an Mg end.  Counted the same way
# on ticks of this workload (profiles/valu_mgh_20x15x15.md, relaxation ticks):
# per active ordered pair for phase 1, per pair with an Mg end for phase 2;
# FP64 ticks, then split ticks as (FP32, FP64).
FLOP_PER_PAIR_TICK = {"phase1_site_sums": 1203, "phase2_gradient": 1944}
FLOP_PER_PAIR_TICK_SPLIT = {"phase1_site_sums": (840, 358), "phase2_gradient": (1768, 144)}
# k_diag (the diagonal Hessian blocks, 3rd-order rule, reference assembly):
# SQ_INSTS_VALU_FLOPS_FP64 x 64 per launch on the same workload, per active
# ordered pair
FLOP_PER_PAIR_DIAG = 12773
PEAK_FP64_VALU = 78.6e12   # MI355X FP64 vector, FLOP/s
PEAK_FP32_VALU = 157.3e12  # MI355X FP32 vector (packed), FLOP/s
HBM_PEAK = 8.0e12          # B/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_mgh_20x15x15_fp64.md")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_mgh_20x15x15.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload():
    import dmdgpu.lattice as lat
    return lat.hcp_mgh(20, 15, 15, x_h=(0.0, 0.1), seed=2024)


def _phase_row(path, phase, col):
    """Cell `col` of the per-phase row `phase` of a committed profile summary."""
    try:
        for line in open(path):
            cells = [c.strip().strip("`") for c in line.split("|")]
            if len(cells) > col + 1 and cells[1] == phase:
                return float(cells[col])
    except (OSError, ValueError):
        pass
    return None


def traffic_per_launch(phase):
    """HBM bytes per launch of a phase's kernels from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel
    factor)."""
    return _phase_row(TRAFFIC_FILE, phase, 2)


def valu_util(phase):
    """VALU issue utilisation of a phase's kernels from the committed rocprofv3
    SQ counter summary (SQ_INSTS_VALU over the issue capacity of their cycles)."""
    return _phase_row(VALU_FILE, phase, 3)


def cpu_baseline(sysd, nnb, nbr, pqe_per_res, budget_s=12.0):
    """The factorised FP64 oracle timed on the host cores (bounded sample:
    whole residuals of the same workload until ~budget_s)."""
    from oracle.oracle import Oracle
    import dmdgpu
    O = Oracle(dmdgpu.TABLE_DIR)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    s = O.system(sysd, nnb, nbr)
    O.fact_residual(s, threads)  # warm
    t0 = time.perf_counter()
    reps = 0
    while True:
        O.fact_residual(s, threads)
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 50:
            break
    dt = time.perf_counter() - t0
    out = {"value": pqe_per_res * reps / dt, "unit": "PQE/s", "cores": threads, "kind": "port",
           "cpu_model": model, "host_cpus": os.cpu_count(),
           "sample": f"{reps} full FP64 residuals of the same 54,000-site workload "
                     f"({dt:.1f} s, factorised oracle, OpenMP)"}
    # The reference's own summation order (O(nb^2) triplets per site), timed on
    # a sample of sites and extrapolated to a whole residual.
    rng = np.random.default_rng(1)
    sites = np.sort(rng.choice(sysd["n"], 128, replace=False)).astype(np.int32)
    mf = O.fact_residual(s, threads)["mf_rho"]
    t0 = time.perf_counter()
    O.ref_residual(s, mf, sites)
    dts = time.perf_counter() - t0
    per_res = dts * sysd["n"] / len(sites)
    out["reference_order"] = {
        "value": pqe_per_res / per_res, "unit": "PQE/s (same PQE count)", "cores": threads,
        "seconds_per_residual": per_res,
        "sample": f"{len(sites)} random sites of the same workload in the reference's O(nb^2) "
                  f"order ({dts:.1f} s), extrapolated to {sysd['n']} sites"}
    # one core, factorised, on the 500-site pure-Mg 5x5x5 crystal
    import dmdgpu.lattice as lat
    c1 = lat.hcp_mgh(5, 5, 5, with_h=False, seed=1)
    st, nn1, nb1 = O.neighbors(c1, dmdgpu.R_CUTOFF)
    s1 = O.system(c1, nn1, nb1)
    pairs1 = int(nn1.sum())
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < 2.0:
        O.fact_residual(s1, 1)
        k += 1
    d1 = time.perf_counter() - t0
    out["one_core_500_sites"] = {"value": 50 * pairs1 * k / d1, "unit": "PQE/s", "cores": 1,
                                 "sample": f"{k} residuals of 500-site hcp Mg 5x5x5 ({d1:.1f} s)"}
    # one whole solve: the 500-site crystal relaxed to |R|_2 <= 1e-8 by the
    # same Newton-CG as the GPU's plain path, on the host cores and on the GPU
    t0 = time.perf_counter()
    ev, rn = cpu_newton_cg(O, c1, nn1, nb1, threads)
    dc = time.perf_counter() - t0
    ctx = dmdgpu.Context(0)
    try:
        ctx.set_state(c1)
        ctx.neighbors()
        opts = dmdgpu.relax_opts(atol=1e-8, max_newton=60, max_cg=100, eta=0.05, precond=0)
        ctx.relax(opts)  # first solve captures the tick graphs
        ctx.set_state(c1)
        ctx.neighbors()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.relax(opts)
        dg = time.perf_counter() - t0
        st = ctx.relax_status(0)
        # the reference's SNES configuration: NGMRES(3), critical-point line search
        ng = dmdgpu.relax_opts(atol=1e-8, max_newton=400, solver="ngmres", ngmres_m=3)
        ctx.set_state(c1)
        ctx.neighbors()
        ctx.relax(ng)
        ctx.set_state(c1)
        ctx.neighbors()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.relax(ng)
        dn = time.perf_counter() - t0
        sn = ctx.relax_status(0)
    finally:
        ctx.close()
    out["c1_solve"] = {"cpu_seconds": dc, "cpu_residual_evals": ev, "cpu_rnorm": rn,
                       "cores": threads, "gpu_seconds": dg, "gpu_residual_evals": st["evals"],
                       "gpu_rnorm": st["rnorm"], "gpu_vs_cpu": dc / dg,
                       "gpu_ngmres3_cp": {"seconds": dn, "residual_evals": sn["evals"],
                                          "iterations": sn["newton"], "rnorm": sn["rnorm"],
                                          "state": sn["state"]},
                       "sample": "500-site hcp Mg 5x5x5 relaxed (positions and widths) to "
                                 "|R|_2 <= 1e-8: Newton-CG, one-sided difference Hessian "
                                 "probes, eta 0.05, no preconditioner, FP64 on both"}
    return out


def cpu_newton_cg(O, sd, nnb, nbr, threads, atol=1e-8, eta=0.05, max_newton=60, max_cg=100,
                  probe=1e-7, cap=0.2):
    """The GPU relaxation's iteration on the oracle's FP64 residual, for the
    CPU solve baseline: unknowns the positions and widths of every site;
    Hessian actions (R(X + e p) - R(X)) / e with RMS displacement `probe`; CG
    to |r| <= eta |R| (steepest step on negative curvature at the first
    iteration); steps capped at `cap` A in position and half of each width;
    backtracking until |R|_2 decreases by 1e-4 of the step.  Returns
    (residual evaluations, final |R|_2)."""
    n = int(sd["n"])
    evals = [0]

    def res(x):
        evals[0] += 1
        d = dict(sd)
        d["q"] = x[:3 * n].reshape(n, 3)
        d["sigma"] = x[3 * n:]
        r = O.fact_residual(O.system(d, nnb, nbr), threads)
        return np.concatenate([r["Rq"].ravel(), r["Rs"]])

    x = np.concatenate([sd["q"].ravel(), sd["sigma"]])
    g = res(x)
    for _ in range(max_newton):
        gn = float(np.linalg.norm(g))
        if gn <= atol:
            break
        d = np.zeros_like(x)
        r = -g
        p = r.copy()
        rr = r @ r
        for it in range(max_cg):
            e = probe * np.sqrt(len(x) / (p @ p))
            ap = (res(x + e * p) - g) / e
            pap = p @ ap
            if pap <= 0:
                if it == 0:
                    d = -g
                break
            a = rr / pap
            d += a * p
            r -= a * ap
            rr1 = r @ r
            if np.sqrt(rr1) <= eta * gn:
                break
            p = r + (rr1 / rr) * p
            rr = rr1
        al = 1.0
        dq = np.abs(d[:3 * n]).max()
        ds = (-d[3 * n:] / x[3 * n:]).max()
        if dq > cap:
            al = cap / dq
        if ds > 0.5:
            al = min(al, 0.5 / ds)
        for _ in range(30):
            gnew = res(x + al * d)
            if np.linalg.norm(gnew) <= (1 - 1e-4 * al) * gn:
                break
            al *= 0.5
        else:
            break
        x = x + al * d
        g = gnew
    return evals[0], float(np.linalg.norm(g))


def split_leg(ctx, args, pqe):
    """The split-precision evaluation (FP32 points, FP64 sums/totals) of the
    same workload: its throughput and its error against the FP64 result."""
    ref = ctx.residual()
    ctx.set_precision("split32")
    cfg = ctx.tuning32
    out = ctx.residual()
    for _ in range(args.warmup):
        ctx.residual_resident()
    ctx.sync()
    ctx.reset_kernel_times()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.residual_resident()
    ctx.sync()
    dt = time.perf_counter() - t0
    kt = ctx.kernel_times()
    ctx.set_timing(False)
    ctx.reset_kernel_times()
    ctx.set_precision("fp64")
    err = {k: float(np.abs(out[k] - ref[k]).max() / np.abs(ref[k]).max()) for k in ("Rq", "Rs", "Rx")}
    err["S0_total"] = float(abs(out["S0_total"] - ref["S0_total"]) / abs(ref["S0_total"]))
    per = {k: v[1] / max(v[0], 1) for k, v in kt.items() if v[0]}
    pairs = ctx.work_counts()[0]
    f32, f64 = (c * pairs for c in FLOP_PER_PAIR_SPLIT["phase2_gradient"])
    t2 = per["phase2_gradient"] / 1e3
    roof = {"bound": "valu", "kernel": "phase2_gradient", "unit": "TFLOP/s",
            "achieved": (f32 + f64) / t2 / 1e12, "achieved_fp32": f32 / t2 / 1e12,
            "achieved_fp64": f64 / t2 / 1e12, "peak_fp32": PEAK_FP32_VALU / 1e12,
            "peak_fp64": PEAK_FP64_VALU / 1e12,
            # time-weighted: each precision's FLOPs at its own published peak
            "frac": (f32 / PEAK_FP32_VALU + f64 / PEAK_FP64_VALU) / t2,
            "avg_launch_ms": per["phase2_gradient"]}
    return {"value": pqe * args.steps / dt, "unit": "PQE/s", "ms_per_step": dt * 1e3 / args.steps,
            "tuning_k1_k2_unroll": list(cfg), "max_rel_err_vs_fp64": err,
            "kernel_ms_per_launch": per, "roofline": roof}


def host_leg(ctx, sysd, pqe, reps=10):
    """The same FP64 residual through the host-buffer boundary (the SNES
    FormFunction form, dmdgpu_residual_packed): X = [q | sigma] in, the packed
    residual [R_q | R_sigma | R_x] out, both in host memory and the caller's
    site order, per call; wall clock, PCIe included.  Also the older form
    with separate arrays (dmdgpu_update_state + dmdgpu_residual)."""
    ctx.set_precision("fp64")
    X = np.concatenate([sysd["q"].ravel(), sysd["sigma"]])
    ctx.residual_packed(X)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.residual_packed(X)
    dt = (time.perf_counter() - t0) / reps
    q, sg = sysd["q"].copy(), sysd["sigma"].copy()
    ctx.update_state(q=q, sigma=sg)
    ctx.residual()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.update_state(q=q, sigma=sg)
        ctx.residual()
    dt2 = (time.perf_counter() - t0) / reps
    # one SNES iteration's device work of the drop-in (the FP64 workload of
    # this configuration): FormFunction + one MatShell product, host buffers
    # and device buffers
    import ctypes as C
    import dmdgpu
    V = np.random.default_rng(1).normal(size=X.size)
    V[3 * sysd["n"]:] *= 0.01
    ctx.jvp_packed(X, V)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.residual_packed(X)
        ctx.jvp_packed(X, V)
    dt3 = (time.perf_counter() - t0) / reps
    dmdgpu.load_library()
    hip = C.CDLL("libamdhip64.so.7")
    ptr = [C.c_void_p() for _ in range(4)]
    sizes = [X.size, X.size, X.size, 5 * sysd["n"]]
    try:
        for p, b in zip(ptr, sizes):
            if hip.hipMalloc(C.byref(p), C.c_size_t(8 * b)):
                raise RuntimeError("hipMalloc failed")
        for p, a in zip(ptr[:2], (X, V)):
            hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * a.size), 1)
        dev = (ptr[0].value, ptr[1].value, ptr[2].value)
        ctx.jvp_packed(device=dev)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.residual_packed(device=(ptr[0].value, ptr[3].value))
            ctx.jvp_packed(device=dev)
        dt4 = (time.perf_counter() - t0) / reps
    finally:
        for p in ptr:
            hip.hipFree(p)
    return {"ms_per_call": dt * 1e3, "value": pqe / dt, "unit": "PQE/s",
            "separate_arrays_ms_per_call": dt2 * 1e3,
            "snes_iteration_residual_plus_jvp_ms": {"host_buffers": dt3 * 1e3,
                                                    "device_buffers": dt4 * 1e3,
                                                    "residual_evaluations": 3}}


def diag_leg(ctx, pairs, reps=5):
    """dmdgpu_diag_blocks on the same workload (reference assembly, FP64): the
    k_diag kernel time from HIP events, the FP64 phase 1 it reads, the wall
    time of the whole call (phase 1 + blocks + copy-out of n x 9 doubles), and
    k_diag's FP64 rate against the vector peak."""
    ctx.set_precision("fp64")
    ctx.diag_blocks()
    ctx.sync()
    ctx.reset_kernel_times()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.diag_blocks()
    dt = (time.perf_counter() - t0) / reps
    ctx.set_timing(False)
    kt = ctx.kernel_times()
    ctx.reset_kernel_times()
    per = lambda k: kt[k][1] / max(kt[k][0], 1)
    ks = per("diag_blocks") * 1e-3
    achieved = FLOP_PER_PAIR_DIAG * pairs / ks
    return {"kernel_ms": ks * 1e3, "phase1_ms": per("phase1_site_sums"), "call_ms": dt * 1e3,
            "roofline": {"bound": "valu", "achieved": achieved / 1e12,
                         "peak": PEAK_FP64_VALU / 1e12, "unit": "TFLOP/s",
                         "frac": achieved / PEAK_FP64_VALU,
                         "flop_per_launch": FLOP_PER_PAIR_DIAG * pairs}}


RULE_POINTS = {3: 12, 5: 27, 44: 44, 6: 6}


def rule5_leg(ctx, args, pqe3, rule=5):
    """The same workload under another pair rule (pair and coincident terms;
    distinct triplets keep the 3rd-order rule): rule 5, the 27-point tensor
    Gauss-Hermite rule, rule 44, the reference's 44-point 6D rule, or rule 6,
    the 6-point Stroud rule on the relative coordinate.  FP64 and
    split-precision time per residual and the split error against FP64; for
    rule 6 also its difference from the reference's 12-point rule (FP64).  PQE
    counts use 13 + (points of the rule) per ordered pair per phase."""
    npt = 13 + RULE_POINTS[rule]
    base = ctx.residual() if rule == 6 else None
    ctx.set_rule(rule)
    out = {}
    res = {}
    try:
        for prec in ("fp64", "split32"):
            ctx.set_precision(prec)
            for _ in range(2):
                ctx.residual_resident()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                ctx.residual_resident()
            ctx.sync()
            dt = (time.perf_counter() - t0) / args.steps
            res[prec] = ctx.residual()
            out[prec] = {"ms_per_step": dt * 1e3, "value": pqe3 / 50 * 2 * npt / dt,
                         "unit": "PQE/s"}
    finally:
        ctx.set_precision("fp64")
        ctx.set_rule(3)
    ref = res["fp64"]
    err = {k: float(np.abs(res["split32"][k] - ref[k]).max() / np.abs(ref[k]).max())
           for k in ("Rq", "Rs", "Rx")}
    err["S0_total"] = float(abs(res["split32"]["S0_total"] - ref["S0_total"]) /
                            abs(ref["S0_total"]))
    out["split32"]["max_rel_err_vs_fp64"] = err
    if base is not None:
        d = {k: float(np.abs(ref[k] - base[k]).max() / np.abs(base[k]).max())
             for k in ("Rq", "Rs", "Rx")}
        d["S0_total"] = float(abs(ref["S0_total"] - base["S0_total"]) / abs(base["S0_total"]))
        out["max_rel_diff_vs_reference_12_point_rule"] = d
    return out


def solve_instances(count):
    """The batched relaxation workload: count/2 Mg-H nanowires along c
    (periodic along the axis, R ~ U[15, 30] A, 10 cells long) and count/2 free
    spheres (R ~ U[12, 25] A), H occupancy x ~ U[0, 0.3], seeded."""
    import dmdgpu.lattice as lat
    return lat.mixed_batch(count, seed=2024)


def _relax_plain(ctx, insts, opts):
    """One instance at a time; returns (seconds, [(status, q, sigma)], per-instance,
    (FP64, FP32) FLOPs)."""
    t = 0.0
    flops = [0.0, 0.0]
    res, per = [], []
    for sd in insts:
        ctx.set_state(sd)
        ctx.neighbors()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.relax(opts)
        ctx.sync()
        dt = time.perf_counter() - t0
        t += dt
        st = ctx.relax_status(0)
        fl = solve_flops(ctx, [0])
        flops[0] += fl[0]
        flops[1] += fl[1]
        q, sg, _ = ctx.instance_state(0)
        res.append((st, q, sg))
        per.append({"sites": int(sd["n"]), "evals": st["evals"],
                    "ms_per_eval": dt * 1e3 / max(st["evals"], 1)})
    ctx.remove_instance(0)
    return t, res, per, flops


def solve_flops(ctx, ids):
    """(FP64, FP32) hardware FLOPs of the residual evaluations of the listed
    instances' solves (relaxation ticks of the Mg lattice with the residual
    merit): per evaluation, phase 1 over every active ordered pair and phase 2
    over those with an Mg end, at the per-pair tick counts of each precision.
    The one-off preconditioner builds are not counted."""
    f64 = f32 = 0.0
    t1, t2 = FLOP_PER_PAIR_TICK["phase1_site_sums"], FLOP_PER_PAIR_TICK["phase2_gradient"]
    s1, s2 = FLOP_PER_PAIR_TICK_SPLIT["phase1_site_sums"], FLOP_PER_PAIR_TICK_SPLIT["phase2_gradient"]
    for i in ids:
        e64, e32 = ctx.relax_evals(i)
        p, pm = ctx.instance_pairs(i)
        f64 += e64 * (t1 * p + t2 * pm)
        f32 += e32 * (s1[0] * p + s2[0] * pm)
        f64 += e32 * (s1[1] * p + s2[1] * pm)
    return f64, f32


def roofline_of(f64, f32, seconds):
    """Time-weighted VALU roofline of a mixed FP64 / FP32 workload."""
    return {"bound": "valu", "unit": "TFLOP/s", "achieved": (f64 + f32) / seconds / 1e12,
            "achieved_fp64": f64 / seconds / 1e12, "achieved_fp32": f32 / seconds / 1e12,
            "peak_fp64": PEAK_FP64_VALU / 1e12, "peak_fp32": PEAK_FP32_VALU / 1e12,
            "frac": (f64 / PEAK_FP64_VALU + f32 / PEAK_FP32_VALU) / seconds,
            "flops_fp64": f64, "flops_fp32": f32}


def _relax_batched(ctx, insts, opts):
    """All instances resident, one launch per tick; returns (seconds, ticks, [(status, q)])."""
    ids = [ctx.add_instance(sd) for sd in insts]
    ctx.sync()
    t0 = time.perf_counter()
    for i in ids:
        ctx.relax_start(i, opts)
    ticks = 0
    live, tick_s = len(ids), []
    while True:
        ta = time.perf_counter()
        running = ctx.relax_tick()
        tick_s.append((live, time.perf_counter() - ta))
        live = running
        if not running:
            break
        ticks += 1
    ctx.sync()
    t = time.perf_counter() - t0
    res = [(ctx.relax_status(i), ctx.instance_state(i)[0]) for i in ids]
    flops = solve_flops(ctx, ids)
    # wall time per tick by the number of solves live in it
    by_live = {}
    for lo, hi in ((1, 4), (5, 16), (17, 48), (49, len(ids))):
        d = [dt for lv, dt in tick_s if lo <= lv <= hi]
        if d:
            by_live[f"{lo}-{hi}"] = {"ticks": len(d), "ms_per_tick": 1e3 * sum(d) / len(d)}
    ev = [ctx.relax_evals(i) for i in ids]
    flops = (flops, by_live, (sum(e[0] for e in ev), sum(e[1] for e in ev)))
    for i in ids:
        ctx.remove_instance(i)
    return t, ticks, res, flops


def solves_leg(tuning, count, graph=True):
    # tuning = ((k1, k2, unroll2) FP64, (k1, k2, unroll2) split)
    """Solves per hour: the plain path (one instance at a time, FP64 Hessian
    probes) against the batched path (all instances resident, one launch per
    tick, split-precision probes, FP64 residuals and stopping test).  Both
    relax Mg positions and widths to |R|_2 <= 1e-8, each with the inner
    tolerance that minimises its own time (CG eta 0.05 for FP64 probes, 0.1
    for split-precision probes with plain CG, 0.05 with the preconditioner).  Run with plain CG and with
    the block-Jacobi preconditioner from the diagonal Hessian blocks (built
    once per solve), the same solver options on both sides of each
    comparison."""
    import dmdgpu
    insts = solve_instances(count)
    sizes = [int(s["n"]) for s in insts]
    n = len(insts)
    out = {"instances": n, "sites_total": sum(sizes), "sites_min": min(sizes),
           "sites_max": max(sizes), "atol": 1e-8}
    ctx = dmdgpu.Context(0)
    t64_plain = None
    if tuning is None:
        # each path tuned on its own work: the batched path with all instances
        # resident, each precision in turn; the plain path's FP64 setting over
        # single instances (the smallest, the median and the largest) timed alone
        from dmdgpu.tune import autotune
        ids = [ctx.add_instance(sd) for sd in insts]
        t32, _ = autotune(ctx, "split32")
        t64, _ = autotune(ctx, "fp64")
        for i in ids:
            ctx.remove_instance(i)
        tuning = (t64, t32)
        order = np.argsort(sizes)
        total = {}
        for k in (order[0], order[len(order) // 2], order[-1]):
            ctx.set_state(insts[k])
            ctx.neighbors()
            for cfg, (p1, p2) in autotune(ctx, "fp64")[1].items():
                a, b = total.get(cfg, (0.0, 0.0))
                total[cfg] = (a + p1, b + p2)
        best1 = min((v[0], k[0]) for k, v in total.items())[1]
        best2 = min((v[1], (k[1], k[2])) for k, v in total.items())[1]
        t64_plain = (best1, best2[0], best2[1])
        log(f"solves: autotuned on the batch: fp64 {t64}, split32 {t32}; "
            f"single instances: fp64 {t64_plain}")
    ctx.set_tuning_precision("fp64", *tuning[0])
    ctx.set_tuning_precision("split32", *tuning[1])
    ctx.set_precision("fp64")
    t64_plain = t64_plain or tuning[0]
    out["tuning_k1_k2_unroll"] = {"fp64": list(tuning[0]), "split32": list(tuning[1]),
                                  "fp64_plain_path": list(t64_plain)}
    ctx.set_graph(graph)
    # preconditioner built once per solve (rebuild interval = max_newton): rebuilding at
    # every Newton step needs the same evaluations on this workload and costs setups
    for pc in (0, 60):
        plain_opts = dmdgpu.relax_opts(atol=1e-8, max_newton=60, max_cg=100, eta=0.05,
                                       mixed=False, precond=pc)
        # split probes: eta 0.1 with plain CG, 0.05 with the preconditioner (sweeps of
        # eta in 0.01..0.3 on this workload)
        mixed_opts = dmdgpu.relax_opts(atol=1e-8, max_newton=60, max_cg=100,
                                       eta=0.05 if pc else 0.1, mixed=True, precond=pc)
        ctx.set_tuning_precision("fp64", *t64_plain)
        t_plain, plain, per_inst, fl_plain = _relax_plain(ctx, insts, plain_opts)
        ctx.set_tuning_precision("fp64", *tuning[0])
        t_batch, ticks, batch, fl_batch = _relax_batched(ctx, insts, mixed_opts)
        dq = max(float(np.abs(b[1] - p[1]).max()) for b, p in zip(batch, plain))
        rec = {
            "plain_fp64_per_instance": {
                "solves_per_hour": n * 3600.0 / t_plain, "seconds": t_plain,
                "converged": sum(p[0]["state"] == "converged" for p in plain),
                "residual_evals": sum(p[0]["evals"] for p in plain),
                "roofline": roofline_of(*fl_plain, t_plain)},
            "batched_split32_probes": {
                "solves_per_hour": n * 3600.0 / t_batch, "seconds": t_batch,
                "converged": sum(b[0]["state"] == "converged" for b in batch),
                "residual_evals": sum(b[0]["evals"] for b in batch),
                "residual_evals_fp64_split": list(fl_batch[2]), "ticks": ticks,
                "ms_per_tick": t_batch * 1e3 / max(ticks, 1),
                "ms_per_tick_by_live_solves": fl_batch[1],
                "roofline": roofline_of(fl_batch[0][0], fl_batch[0][1], t_batch)},
            "speedup": t_plain / t_batch, "max_position_diff_A": dq}
        log(f"plain per instance (precond={pc}): {per_inst}")
        if pc:
            out["block_jacobi_preconditioned"] = rec
        else:
            out.update(rec)
    ctx.close()
    pcr = out["block_jacobi_preconditioned"]
    out["preconditioner_gain"] = {
        "plain": out["plain_fp64_per_instance"]["seconds"] / pcr["plain_fp64_per_instance"]["seconds"],
        "batched": out["batched_split32_probes"]["seconds"] / pcr["batched_split32_probes"]["seconds"]}
    # batched + preconditioned against the plain path without it (different solvers)
    out["batched_preconditioned_vs_plain_unpreconditioned"] = (
        out["plain_fp64_per_instance"]["seconds"] / pcr["batched_split32_probes"]["seconds"])
    return out


def serving_leg(tuning, count, load=0.8, budget_gb=6.0):
    """Serving loop: relaxation and diffusion-step requests (periodic Mg-H boxes
    and nanowires, 500..200,000 sites log-uniform, seeded).  Saturation
    throughput (every request waiting at t = 0) of the plain per-instance FP64
    server and of the continuous-batching server (split-precision probes and
    sweeps, HBM budget admission; also with FP64 probes); then Poisson arrivals at `load` x the plain
    server's capacity, served by each, with p50/p99 latency."""
    import dmdgpu
    import dmdgpu.lattice as lat
    from dmdgpu.serve import Request, latency_stats, poisson_arrivals, serve, serve_plain
    mix = lat.serving_mix(count, seed=7)
    budget = int(budget_gb * 2 ** 30)
    ctx = dmdgpu.Context(0)
    ctx.set_tuning_precision("fp64", *tuning[0])
    ctx.set_tuning_precision("split32", *tuning[1])
    sat = [Request(k, kind, sd, 0.0) for k, (kind, sd) in enumerate(mix)]
    p_sat, t_p = serve_plain(ctx, sat)
    f_sat, t_f = serve(ctx, sat, budget_bytes=budget, mixed=False)
    b_sat, t_b = serve(ctx, sat, budget_bytes=budget, mixed=True)
    rate = load * count / t_p
    arr = poisson_arrivals(count, rate, seed=11)
    pois = [Request(k, kind, sd, float(a)) for k, ((kind, sd), a) in enumerate(zip(mix, arr))]
    p_poi, span_p = serve_plain(ctx, pois)
    b_poi, span_b = serve(ctx, pois, budget_bytes=budget, mixed=True)
    ctx.close()
    dq = max(float(np.abs(p.q - b.q).max()) for p, b in zip(p_sat, b_sat))
    dx = max(float(np.abs(p.x - b.x).max()) for p, b in zip(p_sat, b_sat))
    sizes = [int(sd["n"]) for _, sd in mix]
    return {
        "requests": count, "relax": sum(k == "relax" for k, _ in mix),
        "diffuse": sum(k == "diffuse" for k, _ in mix), "sites_min": min(sizes),
        "sites_max": max(sizes), "sites_total": sum(sizes), "budget_bytes": budget,
        "saturation": {
            "plain_fp64": {"solves_per_hour": count * 3600.0 / t_p, "seconds": t_p,
                           "converged": sum(r.state == "converged" for r in p_sat),
                           "evals": sum(r.evals for r in p_sat)},
            "batched_fp64": {"solves_per_hour": count * 3600.0 / t_f, "seconds": t_f,
                             "converged": sum(r.state == "converged" for r in f_sat),
                             "evals": sum(r.evals for r in f_sat)},
            "batched": {"solves_per_hour": count * 3600.0 / t_b, "seconds": t_b,
                        "converged": sum(r.state == "converged" for r in b_sat),
                        "evals": sum(r.evals for r in b_sat)},
            "speedup": t_p / t_b, "speedup_fp64_batched": t_p / t_f},
        "poisson": {"rate_per_s": rate, "load_vs_plain_capacity": load,
                    "plain_fp64": {"latency_s": latency_stats(p_poi), "makespan_s": span_p},
                    "batched": {"latency_s": latency_stats(b_poi), "makespan_s": span_b}},
        "max_position_diff_A": dq, "max_occupancy_diff": dx}


def copy_calibration(mib, reps=3):
    """Device-to-device copies of `mib` MiB (the HIP runtime's copy kernel):
    under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE they give the counters' scale
    (known bytes read and written per dispatch)."""
    import ctypes as C
    import dmdgpu
    dmdgpu.load_library()  # the HIP runtime the library itself uses
    hip = C.CDLL("libamdhip64.so.7")
    n = C.c_size_t(mib << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), n) or hip.hipMalloc(C.byref(dst), n):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, n):
            raise RuntimeError("hipMemset failed")
        t0 = time.perf_counter()
        for _ in range(reps):
            if hip.hipMemcpy(dst, src, n, 3):  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        dt = (time.perf_counter() - t0) / reps
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    return {"copy_calibration": {"mib": mib, "copies": reps, "bytes_read": mib << 20,
                                 "bytes_written": mib << 20, "ms_per_copy": dt * 1e3}}


def tick_flops_leg(tuning, ticks=4):
    """A few relaxation ticks of the 54,000-site workload (Mg positions and
    widths, residual-norm line search), FP64 probes and then split-precision
    probes, tick kernels launched directly: the run whose hardware FLOP
    counters give the per-pair FLOPs of relaxation ticks (which leave out R_x,
    the pairs with two H ends in phase 2 and the pair-rule values)."""
    import dmdgpu
    ctx = dmdgpu.Context(0)
    try:
        ctx.set_graph(False)
        ctx.set_tuning_precision("fp64", *tuning[0])
        ctx.set_tuning_precision("split32", *tuning[1])
        ctx.set_state(workload())
        ctx.neighbors()
        pairs, pairs_mg = ctx.instance_pairs(0)
        for mixed in (False, True):
            ctx.relax_start(0, dmdgpu.relax_opts(atol=1e-8, max_newton=60, max_cg=100,
                                                 eta=0.05, mixed=mixed))
            for _ in range(ticks):
                ctx.relax_tick()
        ctx.sync()
    finally:
        ctx.close()
    return {"tick_flops": {"pairs": int(pairs), "pairs_mg": int(pairs_mg), "ticks": ticks}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tuning", type=str, default="auto",
                    help="auto, or k1,k2,unroll2[/k1,k2,unroll2] (lanes per site in phase 1/2, "
                         "unroll; FP64 then split precision)")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision leg")
    ap.add_argument("--no-solves", action="store_true", help="skip the solves/hr leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-diag", action="store_true", help="skip the diagonal-Hessian leg")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests in the serving leg")
    ap.add_argument("--instances", type=int, default=64,
                    help="relaxations in the solves/hr leg (half nanowires, half spheres)")
    ap.add_argument("--only-solves", action="store_true",
                    help="run only the solves/hr leg (with --tuning k1,k2,u/k1,k2,u) and print it")
    ap.add_argument("--no-graph", action="store_true",
                    help="solves leg: launch the tick kernels directly instead of as a hipGraph")
    ap.add_argument("--tick-flops", action="store_true",
                    help="only run a few relaxation ticks of the headline workload (FP64, then "
                         "split probes; for the FLOP counters of relaxation ticks) and print "
                         "its pair counts")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device a few times (the HBM counter "
                         "calibration of the traffic profile) and print it")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this benchmark runs on one GPU only; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        print(json.dumps(copy_calibration(args.copy_calibration)), flush=True)
        return
    if args.tick_flops:
        parts = (args.tuning if args.tuning != "auto" else "8,64,0/8,64,1").split("/")
        tun = tuple(tuple(int(v) for v in p.split(",")) for p in (parts[0], parts[-1]))
        print(json.dumps(tick_flops_leg(tun)), flush=True)
        return

    if args.only_solves:
        tun = None
        if args.tuning != "auto":
            parts = args.tuning.split("/")
            tun = tuple(tuple(int(v) for v in p.split(",")) for p in (parts[0], parts[-1]))
        print(json.dumps({"solves": solves_leg(tun, args.instances, graph=not args.no_graph)}),
              flush=True)
        return

    import dmdgpu
    sysd = workload()
    log(f"workload: {sysd['n']} sites")
    ctx = dmdgpu.Context(0)
    ctx.set_state(sysd)
    ctx.neighbors()
    from dmdgpu.tune import autotune
    if args.tuning == "auto":
        cfg32, table32 = autotune(ctx, "split32")
        log(f"autotune split32: {cfg32} {table32}")
        (k1, k2, un), table = autotune(ctx, "fp64")
        log(f"autotune fp64: {(k1, k2, un)} {table}")
    else:
        parts = args.tuning.split("/")
        k1, k2, un = (int(v) for v in parts[0].split(","))
        cfg32 = tuple(int(v) for v in parts[-1].split(","))
        ctx.set_tuning_precision("fp64", k1, k2, un)
        ctx.set_tuning_precision("split32", *cfg32)
        ctx.set_precision("fp64")
    ctx.tuning32 = cfg32
    pairs, pqe = ctx.work_counts()
    log(f"active ordered pairs {pairs}, PQE per residual {pqe}")

    for _ in range(args.warmup):
        ctx.residual_resident()
    ctx.sync()

    ctx.reset_kernel_times()
    ctx.set_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.residual_resident()
    ctx.sync()
    t1 = time.perf_counter()
    ctx.set_timing(False)
    kt = ctx.kernel_times()
    dt = t1 - t0
    log(f"timed {args.steps} steps in {dt:.4f} s; kernels: {kt}")

    # The roofline is that of the largest single kernel, k_phase2 (the pair
    # gradient; its phase adds the slot sums, ~6 % of the phase).  The two
    # phases take about the same time, each over two kernels; both are in the line.
    dom = "phase2_gradient"
    launches, ms = kt[dom]
    avg_s = ms / 1e3 / max(launches, 1)
    achieved = FLOP_PER_PAIR[dom] * pairs / avg_s
    traffic = traffic_per_launch(dom)
    roof = {"bound": "valu", "kernel": dom, "achieved": achieved / 1e12,
            "peak": PEAK_FP64_VALU / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_VALU,
            "traffic": traffic, "avg_launch_ms": avg_s * 1e3,
            "flop_per_launch": FLOP_PER_PAIR[dom] * pairs,
            "hbm_achieved_GBs": traffic / avg_s / 1e9 if traffic else None,
            "hbm_frac": traffic / avg_s / HBM_PEAK if traffic else None,
            "valu_issue_util_profiled": valu_util(dom)}
    l1, ms1 = kt["phase1_site_sums"]
    a1 = ms1 / 1e3 / max(l1, 1)
    roof["phase1_site_sums"] = {
        "achieved": FLOP_PER_PAIR["phase1_site_sums"] * pairs / a1 / 1e12,
        "frac": FLOP_PER_PAIR["phase1_site_sums"] * pairs / a1 / PEAK_FP64_VALU,
        "avg_launch_ms": a1 * 1e3, "traffic": traffic_per_launch("phase1_site_sums"),
        "valu_issue_util_profiled": valu_util("phase1_site_sums")}

    solves = None
    if not args.no_solves:
        solves = solves_leg(None if args.tuning == "auto" else ((k1, k2, un), cfg32),
                            args.instances)
        log(f"solves: {solves}")

    rule5 = rule5_leg(ctx, args, pqe)
    log(f"rule 5: {rule5}")
    rule44 = rule5_leg(ctx, args, pqe, rule=44)
    log(f"rule 44: {rule44}")
    rule6 = rule5_leg(ctx, args, pqe, rule=6)
    log(f"rule 6: {rule6}")

    host = host_leg(ctx, sysd, pqe)
    log(f"host buffers: {host}")

    diag = None
    if not args.no_diag:
        diag = diag_leg(ctx, pairs)
        log(f"diag blocks: {diag}")

    serving = None
    if not args.no_serve:
        serving = serving_leg(((k1, k2, un), cfg32), args.serve_requests)
        log(f"serving: {serving}")

    split = None
    if not args.no_split:
        split = split_leg(ctx, args, pqe)
        log(f"split32: {split}")

    cpu = None
    if not args.no_cpu:
        nn, lst = ctx.get_neighbors()
        cpu = cpu_baseline(sysd, nn, lst, pqe)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "pair-quad evals/s", "value": pqe * args.steps / dt, "unit": "PQE/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Mg-H hcp 20x15x15 periodic: 18,000 Mg + 36,000 tetrahedral H "
                               "(x~U[0,0.1]), reference ADP tables, 3rd-order Stroud rule, "
                               "full FP64 residual per step",
                   "sites": int(sysd["n"]), "active_pairs": int(pairs),
                   "pqe_per_step": int(pqe), "tuning_k1_k2_unroll": [k1, k2, un],
                   "parallelism": "single GPU"},
        "roofline": roof,
        "split32": split,
        "fifth_order_rule": rule5,
        "reference_44_point_rule": rule44,
        "relative_6_point_rule": rule6,
        "host_buffers_pcie_inclusive": host,
        "diag_blocks": diag,
        "solves": solves,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
