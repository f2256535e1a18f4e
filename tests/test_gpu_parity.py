"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: neighbour lists bit-exact; FP64 residual, energies and densities within
1e-10 relative (max-norm of the difference over max-norm of the oracle) -- the
FP64 force-error bar of this path.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, RC

import dmdgpu
import dmdgpu.lattice as lat

pytestmark = pytest.mark.gpu
TOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = {
    "cluster_mgh": lambda: lat.cluster(3, 2, 2, x_h=(0.0, 0.4), seed=3),
    "hcp_mgh_5x3x3": lambda: lat.hcp_mgh(5, 3, 3, x_h=(0.0, 0.3), seed=9),
    "mg_5x5x5_c1": lambda: lat.hcp_mgh(5, 5, 5, with_h=False, seed=1),
    "dump_x5": lambda: lat.read_dump(os.path.join(GOLDEN, "Mg-hcp-cube-x5-x5-x5.dump")),
    # mixed periodicity: periodic along c only, and free on every axis
    "nanowire_r15": lambda: lat.nanowire(15.0, seed=31),
    "sphere_r12": lambda: lat.sphere(12.0, seed=32),
}


def run_both(gpu, oracle, sysd, k1=4, k2=4, unroll2=1, precision="fp64"):
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    assert st == 0
    gpu.set_state(sysd)
    gpu.set_tuning(k1, k2, unroll2)
    gpu.set_precision(precision)
    gpu.neighbors()
    gn, gl = gpu.get_neighbors()
    ref = oracle.fact_residual(oracle.system(sysd, nnb, nbr))
    return (nnb, nbr), (gn, gl), ref


@pytest.mark.parametrize("case", list(CASES))
def test_neighbors_bit_exact(gpu, oracle, case):
    say(f"[gpu] neighbors {case}")
    sysd = CASES[case]()
    (nnb, nbr), (gn, gl), _ = run_both(gpu, oracle, sysd)
    assert np.array_equal(gn, nnb)
    for i in range(sysd["n"]):
        assert np.array_equal(gl[i, :gn[i]], nbr[i, :nnb[i]])


@pytest.mark.parametrize("case", list(CASES))
def test_residual_parity_fp64(gpu, oracle, case):
    say(f"[gpu] residual {case}")
    sysd = CASES[case]()
    _, _, ref = run_both(gpu, oracle, sysd)
    out = gpu.residual()
    assert rel(out["Rq"], ref["Rq"]) < TOL
    assert rel(out["Rs"], ref["Rs"]) < TOL
    assert rel(out["Rx"], ref["Rx"]) < TOL
    assert out["S0_total"] == pytest.approx(ref["S0_total"], rel=TOL)
    assert out["V0_total"] == pytest.approx(ref["V0_total"], rel=TOL)
    assert rel(gpu.mf_rho(), ref["mf_rho"]) < TOL
    V0, S0 = gpu.site_energies()
    assert rel(V0, ref["V0"]) < TOL
    assert rel(S0, ref["S0"]) < TOL


@pytest.mark.parametrize("k1,k2,un", [(4, 16, 1), (16, 4, 0), (8, 8, 1), (4, 4, 0), (8, 32, 0),
                                       (16, 32, 1), (8, 64, 0), (4, 64, 1)])
def test_tuning_variants_agree(gpu, oracle, k1, k2, un):
    say(f"[gpu] tuning {k1} {k2} {un}")
    sysd = CASES["hcp_mgh_5x3x3"]()
    _, _, ref = run_both(gpu, oracle, sysd, k1, k2, un)
    out = gpu.residual()
    assert rel(out["Rq"], ref["Rq"]) < TOL
    assert rel(out["Rs"], ref["Rs"]) < TOL
    assert rel(out["Rx"], ref["Rx"]) < TOL
    gpu.set_tuning(4, 4, 1)


def test_tuning_rejects_unsupported_values(gpu):
    """k_phase1 in {4, 8, 16}, k_phase2 in {4, 8, 16, 32, 64}; anything else is an
    argument error and leaves the tuning unchanged."""
    import dmdgpu
    for bad in ((32, 8, 0), (8, 128, 0), (2, 8, 0), (8, 12, 1)):
        with pytest.raises(dmdgpu.DmdError):
            gpu.set_tuning(*bad)
    gpu.set_tuning(4, 4, 1)


# Split precision (FP32 point evaluation, FP64 distances/sums/totals) against
# the FP64 oracle.  Stated per quantity: forces R_q and width residuals R_sigma
# within 1e-6 relative (the accuracy bar of the split path), occupancy
# residuals and densities within 1e-7, totals within 1e-8.
SPLIT_TOL = {"Rq": 1e-6, "Rs": 1e-6, "Rx": 1e-7, "mf_rho": 1e-7, "S0_total": 1e-8}


@pytest.mark.parametrize("case", ["hcp_mgh_5x3x3", "mg_5x5x5_c1", "cluster_mgh", "nanowire_r15",
                                  "sphere_r12"])
@pytest.mark.parametrize("tun", [(4, 4, 0), (4, 4, 1), (8, 64, 1), (8, 32, 0)])
def test_split32_against_fp64(gpu, oracle, case, tun):
    """Split precision against the FP64 oracle at the autotuner's candidates
    (the bench picks 8, 64, 1 for the split path)."""
    say(f"[gpu] split32 {case} tuning={tun}")
    sysd = CASES[case]()
    _, _, ref = run_both(gpu, oracle, sysd, *tun, "split32")
    out = gpu.residual()
    err = {k: rel(out[k], ref[k]) for k in ("Rq", "Rs", "Rx")}
    err["mf_rho"] = rel(gpu.mf_rho(), ref["mf_rho"])
    err["S0_total"] = abs(out["S0_total"] - ref["S0_total"]) / abs(ref["S0_total"])
    say(f"    split32 errors {err}")
    gpu.set_precision("fp64")
    gpu.set_tuning(4, 4, 1)
    for k, tol in SPLIT_TOL.items():
        assert err[k] < tol, (k, err[k])


def test_reference_order_spot_check(gpu, oracle):
    """GPU against the reference-order restatement directly (small cluster)."""
    say("[gpu] reference-order spot check")
    sysd = lat.cluster(2, 1, 1, x_h=(0.0, 0.4), seed=3)
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    s = oracle.system(sysd, nnb, nbr)
    rho = oracle.ref_mf_rho(s)
    Rq, Rs, Rx = oracle.ref_residual(s, rho)
    gpu.set_state(sysd)
    gpu.neighbors()
    out = gpu.residual()
    assert rel(gpu.mf_rho(), rho) < TOL
    assert rel(out["Rq"], Rq) < TOL
    assert rel(out["Rs"], Rs) < TOL
    h = sysd["species"] == 1
    assert rel(out["Rx"][h], Rx[h]) < TOL


def test_update_state_and_determinism(gpu, oracle):
    say("[gpu] update_state + determinism")
    sysd = CASES["hcp_mgh_5x3x3"]()
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    gpu.set_state(sysd)
    gpu.neighbors()
    rng = np.random.default_rng(5)
    sysd["q"] = sysd["q"] + rng.normal(0, 0.02, sysd["q"].shape)
    sysd["sigma"] = sysd["sigma"] * (1 + 0.05 * rng.random(sysd["n"]))
    hs = sysd["species"] == 1
    sysd["x"][hs] = rng.uniform(0, 0.3, hs.sum())
    gpu.update_state(sysd["q"], sysd["sigma"], sysd["x"])
    a = gpu.residual()
    b = gpu.residual()
    for k in ("Rq", "Rs", "Rx"):
        assert np.array_equal(a[k], b[k])
    ref = oracle.fact_residual(oracle.system(sysd, nnb, nbr))
    assert rel(a["Rq"], ref["Rq"]) < TOL
    assert rel(a["Rx"], ref["Rx"]) < TOL


def test_full_size_mgh_20x15x15(gpu, oracle):
    """The 54,000-site Mg-H workload (18k Mg + 36k tetrahedral H):
    neighbour lists bit-exact and every residual entry within 1e-10."""
    say("[gpu] full-size 20x15x15")
    sysd = lat.hcp_mgh(20, 15, 15, x_h=(0.0, 0.1), seed=2024)
    assert sysd["n"] == 54000
    (nnb, nbr), (gn, gl), ref = run_both(gpu, oracle, sysd)
    assert np.array_equal(gn, nnb)
    assert np.array_equal(np.where(np.arange(250)[None, :] < gn[:, None], gl, -1),
                          np.where(np.arange(250)[None, :] < nnb[:, None], nbr, -1))
    out = gpu.residual()
    assert rel(out["Rq"], ref["Rq"]) < TOL
    assert rel(out["Rs"], ref["Rs"]) < TOL
    assert rel(out["Rx"], ref["Rx"]) < TOL
    assert out["S0_total"] == pytest.approx(ref["S0_total"], rel=TOL)
    # size-independent property: translation invariance of the periodic sum
    assert np.abs(out["Rq"].sum(0)).max() < 1e-9 * np.abs(out["Rq"]).max() * sysd["n"]


@pytest.mark.parametrize("case", ["hcp_mgh_5x3x3", "cluster_mgh"])
def test_zero_temperature_density(gpu, oracle, case):
    """The zero-temperature density (reference evaluate_rho_i_adp_MgHx,
    src/ADP/MgHx-mf-V-bulk.cpp:83-143) as the sigma -> 0 limit of the GPU
    mean-field density: at sigma = 1e-9 A the rule's points sit within 3e-9 A
    of r_ij and, being symmetric, cancel the first-order change, so dmdgpu_mf_rho
    equals the oracle's zero-T density to rounding."""
    say(f"[gpu] zero-T density {case}")
    sysd = CASES[case]()
    sysd["sigma"][:] = 1e-9
    (nnb, nbr), _, _ = run_both(gpu, oracle, sysd)
    rho = gpu.mf_rho()
    ref = oracle.ref_rho0(oracle.system(sysd, nnb, nbr))
    assert rel(rho, ref) < 1e-12


def test_error_codes(gpu):
    say("[gpu] error codes")
    small = lat.hcp_mgh(3, 2, 2, seed=0)  # periodic edges < 2 r_c
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.set_state(small)
    assert e.value.code == 7
    bad = lat.cluster(2, 1, 1, seed=0)
    bad["species"][0] = 13
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.set_state(bad)
    assert e.value.code == 4
    iso = lat.cluster(2, 1, 1, seed=0)
    iso["q"][0] += 100.0
    gpu.set_state(iso)
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.neighbors()
    assert e.value.code == 6
    dense = lat.cluster(6, 4, 4, a=1.2, seed=0)  # 1152 sites, most within r_c
    gpu.set_state(dense)
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.neighbors()
    assert e.value.code == 5
    empty = {k: (v[:0] if isinstance(v, np.ndarray) and v.ndim and k not in
                 ("lo", "hi", "periodic") else v) for k, v in small.items()}
    empty["n"] = 0
    with pytest.raises(dmdgpu.DmdError) as e:
        gpu.add_instance(empty)  # an instance needs at least one site
    assert e.value.code == 1


# Other pair rules for the pair and coincident terms -- 5: 27-point tensor
# Gauss-Hermite on the relative coordinate; 44: the reference's 44-point 6D
# rule; 6: the 6-point Stroud rule on the relative coordinate -- FP64 GPU
# against the oracle's restatement of the same rule at the FP64 bar, and the
# split evaluation against FP64 at the split bars.
@pytest.mark.parametrize("rule", [5, 44, 6])
@pytest.mark.parametrize("case", ["hcp_mgh_5x3x3", "cluster_mgh"])
def test_fifth_order_rule(gpu, oracle, case, rule):
    say(f"[gpu] rule {rule} {case}")
    sysd = CASES[case]()
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    ref = oracle.fact_residual(oracle.system(sysd, nnb, nbr, rule))
    gpu.set_state(sysd)
    gpu.set_tuning(8, 8, 0)
    gpu.neighbors()
    gpu.set_rule(rule)
    try:
        gpu.set_precision("fp64")
        out = gpu.residual()
        for k in ("Rq", "Rs", "Rx"):
            assert rel(out[k], ref[k]) < TOL, k
        assert out["S0_total"] == pytest.approx(ref["S0_total"], rel=TOL)
        assert rel(gpu.mf_rho(), ref["mf_rho"]) < TOL
        gpu.set_precision("split32")
        out = gpu.residual()
        err = {k: rel(out[k], ref[k]) for k in ("Rq", "Rs", "Rx")}
        err["mf_rho"] = rel(gpu.mf_rho(), ref["mf_rho"])
        err["S0_total"] = abs(out["S0_total"] - ref["S0_total"]) / abs(ref["S0_total"])
        say(f"    rule-{rule} split32 errors {err}")
        for k, tol in SPLIT_TOL.items():
            assert err[k] < tol, (k, err[k])
    finally:
        gpu.set_precision("fp64")
        gpu.set_rule(3)
        gpu.set_tuning(4, 4, 1)
