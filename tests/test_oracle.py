"""CPU tests that pin the oracle (no GPU).

The reference has no tests or golden vectors for this path, so the oracle is
pinned by: the reference's own table data and its closed-form coefficient
relations; closed forms of the quadrature rule on polynomial tables; an
independent numpy enumeration of the reference's quadrature points; agreement of
the two oracle forms (reference order vs factorised); finite differences of the
free-entropy total; and symmetry (translation invariance, zero forces on a
perfect hcp crystal).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, RC, TABLE_DIR

import dmdgpu.lattice as lat
import dmdgpu.tables as tabs
from oracle.oracle import Oracle


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ---------------------------------------------------------------------------
# Tables: src/ADP/MgHx-ADP.cpp:366-501, src/Numerical/cubic-spline.cpp:84-128
# ---------------------------------------------------------------------------

def _raw_blocks(path, alloy):
    vals = open(path).read().split()
    pos = 1 if alloy else 3
    blocks = []
    while pos < len(vals):
        n, dx = int(vals[pos]), float(vals[pos + 1])
        rows = np.array(vals[pos + 2:pos + 2 + 9 * n], dtype=np.float64).reshape(n, 9)
        blocks.append((n, dx, rows))
        pos += 2 + 9 * n
    return blocks


@pytest.mark.parametrize("fname,alloy", [("adpmg.dat", False), ("adph.dat", False),
                                         ("adpmgh.dat", True)])
def test_reference_table_columns(fname, alloy):
    """Derived columns satisfy db=b, dc=2c, dd=3d, ddc=2c, ddd=6d."""
    blocks = _raw_blocks(os.path.join(TABLE_DIR, fname), alloy)
    assert len(blocks) == (3 if alloy else 5)
    for n, dx, r in blocks:
        a, b, c, d, db, dc, dd, ddc, ddd = r.T
        scale = np.abs(r).max()
        for lhs, rhs in ((db, b), (dc, 2 * c), (dd, 3 * d), (ddc, 2 * c), (ddd, 6 * d)):
            assert np.abs(lhs - rhs).max() <= 1e-9 * scale


def test_reference_table_shapes(oracle):
    # inputs/adpmg.dat:1-2,1197 -- 1194 r-segments of 0.00527527..., 967 embedding segments
    assert oracle.spline_size(0, 0) == (1194, pytest.approx(0.00527527527528))
    assert oracle.spline_size(1, 0) == (967, pytest.approx(0.000982083660807))
    for which in (2, 3, 4):
        for slot in (0, 1, 2):
            assert oracle.spline_size(which, slot)[0] == 1194


def test_spline_clamp_and_horner(oracle):
    """Value/derivative use the Horner forms; past n dx the last segment's end
    value and end slope are returned (cubic-spline.cpp:89-93)."""
    blocks = _raw_blocks(os.path.join(TABLE_DIR, "adpmg.dat"), False)
    n, dx, rows = blocks[3]  # u
    for x in (0.7, 2.3456, 5.0):
        m = int(x / dx)
        p = x - m * dx
        a, b, c, d, db, dc, dd, ddc, ddd = rows[m]
        assert oracle.spline(3, 0, 0, x) == pytest.approx(a + (b + (c + d * p) * p) * p, rel=1e-14)
        assert oracle.spline(3, 0, 1, x) == pytest.approx(db + (dc + dd * p) * p, rel=1e-14)
    a, b, c, d, db, dc, dd, ddc, ddd = rows[-1]
    end_v = a + (b + (c + d * dx) * dx) * dx
    end_d = db + (dc + dd * dx) * dx
    for x in (n * dx + 1e-9, n * dx + 0.3, 7.5):
        assert oracle.spline(3, 0, 0, x) == pytest.approx(end_v, rel=1e-14, abs=1e-300)
        assert oracle.spline(3, 0, 1, x) == pytest.approx(end_d, rel=1e-14, abs=1e-300)


# ---------------------------------------------------------------------------
# Neighbours: src/Atoms/Neighbors.cpp:39-134
# ---------------------------------------------------------------------------

def numpy_neighbors(sysd, rc):
    q = sysd["q"]
    L = sysd["hi"] - sysd["lo"]
    d = q[:, None, :] - q[None, :, :]
    for a in range(3):
        if sysd["periodic"][a]:
            d[..., a] -= L[a] * np.rint(d[..., a] / L[a])
    r = np.sqrt((d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2)
    np.fill_diagonal(r, np.inf)
    return [np.nonzero(r[i] <= rc)[0] for i in range(len(q))]


@pytest.mark.parametrize("maker", [lambda: lat.hcp_mgh(5, 3, 3, seed=1),
                                   lambda: lat.cluster(3, 2, 2, seed=2),
                                   lambda: lat.read_dump(os.path.join(GOLDEN, "Mg-hcp-cube-x5-x5-x5.dump"))])
def test_neighbors_match_numpy(oracle, maker):
    sysd = maker()
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    assert st == 0
    ref = numpy_neighbors(sysd, RC)
    for i in range(sysd["n"]):
        assert list(nbr[i, :nnb[i]]) == list(ref[i])


def test_dump_fixture_reader():
    d = lat.read_dump(os.path.join(GOLDEN, "Mg-hcp-cube-x5-x5-x5.dump"))
    assert d["n"] == 500 and (d["species"] == 12).all() and list(d["periodic"]) == [1, 1, 1]
    assert d["beta"][0] == pytest.approx(38.6817305917)
    assert d["sigma"][0] == pytest.approx(0.01)


def test_dump_writer_round_trip(tmp_path):
    """write_dump -> read_dump is exact, for the reference fixture and for a
    generated nanowire with mixed boundary flags."""
    src = lat.read_dump(os.path.join(GOLDEN, "Mg-hcp-cube-x5-x5-x5.dump"))
    for sd in (src, lat.nanowire(15, seed=3)):
        path = tmp_path / "out.dump"
        lat.write_dump(path, sd, step=7)
        back = lat.read_dump(path)
        assert back["n"] == sd["n"]
        for k in ("q", "sigma", "x", "beta", "gamma", "lo", "hi"):
            assert np.array_equal(back[k], sd[k]), k
        assert np.array_equal(back["species"], sd["species"])
        assert list(back["periodic"]) == list(sd["periodic"])
        assert open(path).read().splitlines()[1] == "7"


def test_neighbor_errors(oracle):
    sysd = lat.cluster(2, 1, 1, seed=0)
    sysd["q"][0] += 100.0  # isolated site
    st, _, _ = oracle.neighbors(sysd, RC)
    assert st == 2


# ---------------------------------------------------------------------------
# Quadrature rule closed forms on polynomial tables
# ---------------------------------------------------------------------------

def _poly_tables(tmp, rho, F, phi, u, w):
    fn = {"rho_Mg": tabs.poly(*rho), "rho_H": tabs.poly(*rho), "F_Mg": tabs.poly(*F),
          "F_H": tabs.poly(*F)}
    for pp in ("MgMg", "HH", "MgH"):
        fn["phi_" + pp] = tabs.poly(*phi)
        fn["u_" + pp] = tabs.poly(*u)
        fn["w_" + pp] = tabs.poly(*w)
    tabs.write_tables(str(tmp), fn, nr=2400, dxr=0.005, ne=2000, dxe=0.05)
    return Oracle(str(tmp))


def _small_cluster(seed=4, with_h=True):
    s = lat.cluster(2, 1, 1, with_h=with_h, x_h=(0.2, 0.9), seed=seed, sigma_jitter=0.3)
    return s


def test_quadrature_closed_form_density_and_dipole(tmp_path):
    """rho = 0.01 r^2, F = rho_bar, u = 1, phi = w = 0:
    rho_bar_i = 0.01 sum_j n_j (|r_ij|^2 + 3 (s_i^2 + s_j^2))   (6D rule exact to degree 3)
    V0_i = n_i rho_bar_i + 1/2 n_i^2 [ |sum_j n_j r_ij|^2 + 3 s_i^2 (sum_j n_j)^2
                                       + 3 sum_j n_j^2 s_j^2 ]"""
    O = _poly_tables(tmp_path, rho=(0, 0, 0.01), F=(0, 1), phi=(0,), u=(1,), w=(0,))
    sysd = _small_cluster()
    st, nnb, nbr = O.neighbors(sysd, RC)
    assert st == 0
    s = O.system(sysd, nnb, nbr)
    f = O.fact_residual(s)
    rho_ref = O.ref_mf_rho(s)
    V0_ref = O.ref_V0(s, rho_ref)
    q, sg, x = sysd["q"], sysd["sigma"], sysd["x"]
    rho_cf = np.zeros(sysd["n"])
    V0_cf = np.zeros(sysd["n"])
    for i in range(sysd["n"]):
        js = nbr[i, :nnb[i]]
        r = q[i] - q[js]
        nj = x[js]
        rho_cf[i] = 0.01 * np.sum(nj * ((r ** 2).sum(1) + 3 * (sg[i] ** 2 + sg[js] ** 2)))
        G = (nj[:, None] * r).sum(0)
        V0_cf[i] = x[i] * rho_cf[i] + 0.5 * x[i] ** 2 * (
            G @ G + 3 * sg[i] ** 2 * nj.sum() ** 2 + 3 * np.sum(nj ** 2 * sg[js] ** 2))
    assert rel(rho_ref, rho_cf) < 1e-12
    assert rel(f["mf_rho"], rho_cf) < 1e-12
    assert rel(V0_ref, V0_cf) < 1e-12
    assert rel(f["V0"], V0_cf) < 1e-12


def test_fifth_order_pair_rule_closed_form(tmp_path):
    """Coincident quadrupole term of a dimer with w = c (all other functions
    zero): V0_i = n_i^2 n_j^2 c^2/3 E|r|^4 over r ~ N(r_ij, s^2 I),
    s^2 = sigma_i^2 + sigma_j^2, whose Gaussian moment is
    |r_ij|^4 + 10 |r_ij|^2 s^2 + 15 s^4.  The 27-point tensor rule (degree 5)
    is exact, and so is the reference's 44-point 6D degree-5 rule (rule 44);
    the reference's 12-point rule (degree 3) gives 18 (sigma_i^4 + sigma_j^4)
    in place of 15 s^4, and the 6-point relative rule (rule 6, degree 3)
    9 s^4 -- (|r|^2 + 3 s^2)^2 + 4 |r|^2 s^2 over its points -- which the
    test also pins."""
    c = 0.7
    O = _poly_tables(tmp_path, rho=(0,), F=(0,), phi=(0,), u=(0,), w=(c,))
    sysd = lat.cluster(1, 1, 1, with_h=False, seed=3, sigma_jitter=0.3)
    keep = [0, 1]
    for k in ("q", "sigma", "x", "species", "beta", "gamma"):
        sysd[k] = sysd[k][keep]
    sysd["n"] = 2
    sysd["x"] = np.array([0.9, 0.6])
    st, nnb, nbr = O.neighbors(sysd, RC)
    assert st == 0 and list(nnb) == [1, 1]
    A = ((sysd["q"][0] - sysd["q"][1]) ** 2).sum()
    si, sj = sysd["sigma"]
    s2 = si ** 2 + sj ** 2
    pre = (sysd["x"][0] * sysd["x"][1]) ** 2 * c * c / 3.0
    exact = pre * (A * A + 10 * A * s2 + 15 * s2 * s2)
    third = pre * (A * A + 10 * A * s2 + 18 * (si ** 4 + sj ** 4))
    six = pre * (A * A + 10 * A * s2 + 9 * s2 * s2)
    for rule, want in ((5, exact), (44, exact), (3, third), (6, six)):
        V0 = O.fact_residual(O.system(sysd, nnb, nbr, rule))["V0"]
        assert V0 == pytest.approx([want, want], rel=1e-12)


def test_relative_six_point_rule_equals_reference_rule_at_equal_widths(oracle):
    """With sigma_i = sigma_j on every pair the 6-point relative rule puts
    the same points (r +- sqrt(6) sigma e_l) with the same weights as the
    reference's 12-point rule, and its width derivatives agree too (sqrt(3/2)
    on six points of weight 1/6 against sqrt(6) on the i-half of weight 1/12):
    every output equal to rounding.  With unequal widths they differ."""
    sysd = lat.hcp_mgh(4, 3, 3, x_h=(0.05, 0.5), seed=21)
    sysd["sigma"][:] = 0.12
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    a = oracle.fact_residual(oracle.system(sysd, nnb, nbr, 3))
    b = oracle.fact_residual(oracle.system(sysd, nnb, nbr, 6))
    for k in ("Rq", "Rs", "Rx", "V0", "mf_rho"):
        assert np.abs(a[k] - b[k]).max() <= 1e-12 * np.abs(a[k]).max(), k
    sysd = lat.hcp_mgh(4, 3, 3, x_h=(0.05, 0.5), seed=21, sigma_jitter=0.2)
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    a = oracle.fact_residual(oracle.system(sysd, nnb, nbr, 3))
    b = oracle.fact_residual(oracle.system(sysd, nnb, nbr, 6))
    assert np.abs(a["V0"] - b["V0"]).max() > 1e-9 * np.abs(a["V0"]).max()


def numpy_stroud_V0(sysd, nnb, nbr, rho, F, phi, u, w):
    """Independent enumeration of the reference's rule for V0_i: pair integrals
    with sqrt(6) sigma shifts (weight 1/12), distinct triplets with 3 sigma
    shifts of one site (weight 1/18, factor 2 for j1 < j2), coincident triplets
    with i or j (both copies) shifted by sqrt(6) sigma (weight 1/12).
    rho/F/phi/u/w are plain numpy callables (polynomials)."""
    q, sg, x = sysd["q"], sysd["sigma"], sysd["x"]
    E = np.eye(3)

    def dip_quad(qi, q1, q2, ni, n1, n2):
        r1, r2 = qi - q1, qi - q2
        l1, l2 = np.linalg.norm(r1), np.linalg.norm(r2)
        a1, a2 = ni * n1 * u(l1), ni * n2 * u(l2)
        b1, b2 = ni * n1 * w(l1), ni * n2 * w(l2)
        dot = r1 @ r2
        return 0.5 * a1 * a2 * dot + 0.5 * b1 * b2 * dot ** 2 - b1 * b2 * (r1 @ r1) * (r2 @ r2) / 6

    V0 = np.zeros(sysd["n"])
    for i in range(sysd["n"]):
        js = nbr[i, :nnb[i]]
        rb, vp, vt = 0.0, 0.0, 0.0
        for j in js:
            acc_r, acc_p = 0.0, 0.0
            for who in (0, 1):
                for l in range(3):
                    for s in (1, -1):
                        d = s * np.sqrt(6) * (sg[i] if who == 0 else sg[j]) * E[l]
                        qi = q[i] + (d if who == 0 else 0)
                        qj = q[j] + (d if who == 1 else 0)
                        rr = np.linalg.norm(qi - qj)
                        acc_r += x[j] * rho(rr) / 12
                        acc_p += 0.5 * x[i] * x[j] * phi(rr) / 12
            rb += acc_r
            vp += acc_p
        for a, j1 in enumerate(js):
            for j2 in js[a:]:
                if j1 == j2:
                    acc = 0.0
                    for who in (0, 1):
                        for l in range(3):
                            for s in (1, -1):
                                d = s * np.sqrt(6) * (sg[i] if who == 0 else sg[j1]) * E[l]
                                qi = q[i] + (d if who == 0 else 0)
                                qj = q[j1] + (d if who == 1 else 0)
                                acc += dip_quad(qi, qj, qj, x[i], x[j1], x[j1]) / 12
                    vt += acc
                else:
                    acc = 0.0
                    for who in range(3):
                        site = (i, j1, j2)[who]
                        for l in range(3):
                            for s in (1, -1):
                                pos = [q[i].copy(), q[j1].copy(), q[j2].copy()]
                                pos[who] = pos[who] + s * 3.0 * sg[site] * E[l]
                                acc += dip_quad(*pos, x[i], x[j1], x[j2]) / 18
                    vt += 2 * acc
        V0[i] = x[i] * F(rb) + vp + vt
    return V0


def test_numpy_enumeration_of_stroud_points(tmp_path):
    """Both oracle forms equal an independent enumeration of the reference's
    quadrature points for polynomial u, w, phi, rho, F (all terms on)."""
    c = dict(rho=(0.02, -0.001, 0.002), F=(0.1, -0.5, 0.3), phi=(0.4, -0.1, 0.004, 0.0002),
             u=(0.3, -0.02, 0.001), w=(0.1, 0.01, -0.001))
    O = _poly_tables(tmp_path, **c)
    sysd = _small_cluster(seed=7)
    st, nnb, nbr = O.neighbors(sysd, RC)
    s = O.system(sysd, nnb, nbr)
    f = O.fact_residual(s)
    V0_ref = O.ref_V0(s, O.ref_mf_rho(s))
    P = {k: np.polynomial.polynomial.Polynomial(v) for k, v in c.items()}
    V0_np = numpy_stroud_V0(sysd, nnb, nbr, **P)
    assert rel(V0_ref, V0_np) < 1e-12
    assert rel(f["V0"], V0_np) < 1e-12


# ---------------------------------------------------------------------------
# Reference order vs factorised, with the reference tables
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("maker", [lambda: lat.cluster(2, 1, 1, x_h=(0.0, 0.4), seed=3),
                                   lambda: lat.hcp_mgh(4, 3, 3, x_h=(0.0, 0.3), seed=9)])
def test_reference_order_equals_factorised(oracle, maker):
    sysd = maker()
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    assert st == 0
    s = oracle.system(sysd, nnb, nbr)
    f = oracle.fact_residual(s)
    rho_ref = oracle.ref_mf_rho(s)
    assert rel(f["mf_rho"], rho_ref) < 1e-13
    n = sysd["n"]
    sites = np.arange(n) if n < 100 else np.random.default_rng(0).choice(n, 12, replace=False)
    assert rel(f["V0"][sites], oracle.ref_V0(s, rho_ref, sites)) < 1e-13
    assert rel(f["S0"][sites], oracle.ref_S0(s, rho_ref, sites)) < 1e-13
    Rq, Rs, Rx = oracle.ref_residual(s, rho_ref, sites)
    assert rel(f["Rq"][sites], Rq) < 1e-12
    assert rel(f["Rs"][sites], Rs) < 1e-12
    h = sysd["species"][sites] == 1
    # Mg occupancy is not an unknown; the reference's Mg embedding term for
    # i == u omits F(rho_bar) (MgHx-mf-V-bulk.cpp:2311-2315), so compare H rows.
    assert rel(f["Rx"][sites][h], Rx[h]) < 1e-12


def test_masked_sites_and_zero_occupancy(oracle):
    sysd = lat.cluster(2, 1, 1, x_h=(0.0, 0.3), seed=11)
    hs = np.nonzero(sysd["species"] == 1)[0]
    sysd["x"][hs[:4]] = [0.0, 1e-4, 4.99e-4, 5.01e-4]
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    s = oracle.system(sysd, nnb, nbr)
    f = oracle.fact_residual(s)
    for k in range(3):
        u = hs[k]
        assert f["V0"][u] == 0 and f["S0"][u] == 0 and (f["Rq"][u] == 0).all()
    rho_ref = oracle.ref_mf_rho(s)
    assert rel(f["mf_rho"], rho_ref) < 1e-13
    Rq, Rs, _ = oracle.ref_residual(s, rho_ref)
    assert rel(f["Rq"], Rq) < 1e-12 and rel(f["Rs"], Rs) < 1e-12


# ---------------------------------------------------------------------------
# Finite differences and symmetry
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def synth(tmp_path_factory):
    d = tmp_path_factory.mktemp("synth")
    tabs.synthetic_mgh(str(d), seed=3, nr=3000, ne=3000)
    return Oracle(str(d))


@pytest.mark.parametrize("rule", [3, 5, 44, 6])
def test_finite_differences_of_free_entropy(synth, rule):
    """R = d(sum_i S0_i)/d(q, sigma, x) on smooth tables (consistent clamp)."""
    sysd = lat.hcp_mgh(4, 3, 3, x_h=(0.05, 0.5), seed=21)
    st, nnb, nbr = synth.neighbors(sysd, 6.5)
    assert st == 0
    base = synth.fact_residual(synth.system(sysd, nnb, nbr, rule))

    def tot(key, u, a, h):
        sd = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sysd.items()}
        if key == "q":
            sd["q"][u, a] += h
        else:
            sd[key][u] += h
        return synth.fact_residual(synth.system(sd, nnb, nbr, rule))["S0_total"]

    rng = np.random.default_rng(1)
    for u in rng.choice(sysd["n"], 4, replace=False):
        for a in range(3):
            h = 1e-4
            fd = (tot("q", u, a, h) - tot("q", u, a, -h)) / (2 * h)
            assert fd == pytest.approx(base["Rq"][u, a], rel=1e-5, abs=1e-5)
        h = 1e-5
        fd = (tot("sigma", u, 0, h) - tot("sigma", u, 0, -h)) / (2 * h)
        assert fd == pytest.approx(base["Rs"][u], rel=1e-6)
        if sysd["species"][u] == 1:
            h = 1e-6
            fd = (tot("x", u, 0, h) - tot("x", u, 0, -h)) / (2 * h)
            assert fd == pytest.approx(base["Rx"][u], rel=1e-6)


def test_per_pair_derivatives_by_finite_differences(synth):
    """The reference-order per-(u, i) derivatives (evaluate_DV0_i_Dq_u_MgHx,
    src/ADP/MgHx-mf-V-bulk.cpp:1823-2098, and evaluate_DV0_i_Dxi_u_adp_MgHx,
    :2102-2325) against central differences of the reference-order V0_i with
    the densities re-evaluated, one (u, i) at a time: u = i and two neighbours
    u of an Mg centre and of an H centre (x only for H centres, where the
    reference's form is the exact derivative)."""
    sysd = lat.cluster(2, 2, 2, x_h=(0.05, 0.5), seed=4)
    st, nnb, nbr = synth.neighbors(sysd, 6.5)
    assert st == 0
    s0 = synth.system(sysd, nnb, nbr)
    rho0 = synth.ref_mf_rho(s0)

    def v0(i, key, u, a, h):
        sd = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sysd.items()}
        if key == "q":
            sd["q"][u, a] += h
        else:
            sd[key][u] += h
        s = synth.system(sd, nnb, nbr)
        return synth.ref_V0(s, synth.ref_mf_rho(s), [i])[0]

    def fd(i, key, u, a, h):
        return (v0(i, key, u, a, h) - v0(i, key, u, a, -h)) / (2 * h)

    rng = np.random.default_rng(2)
    act = sysd["x"] >= 5e-4
    centres = [int(rng.choice(np.flatnonzero(sysd["species"] == lat.MG))),
               int(rng.choice(np.flatnonzero((sysd["species"] == lat.H) & act)))]
    checked = 0
    for i in centres:
        row = [int(v) for v in nbr[i, :nnb[i]] if act[v]]
        for u in [i, row[0], row[len(row) // 2]]:
            d = synth.ref_DV0_Dq(s0, u, i, rho0)
            scale = np.abs(d).max() * 1e-7 + 1e-9  # FD rounding: |V0| eps / h ~ 1e-11
            for a in range(3):
                assert fd(i, "q", u, a, 1e-5) == pytest.approx(d[a], rel=1e-6, abs=scale)
            assert fd(i, "sigma", u, 0, 1e-6) == pytest.approx(d[3], rel=1e-6, abs=scale)
            if sysd["species"][i] == lat.H:
                dx = synth.ref_DV0_Dxi(s0, u, i, rho0)
                assert fd(i, "x", u, 0, 1e-6) == pytest.approx(dx, rel=1e-6, abs=1e-9)
            checked += 1
    assert checked == 6


def test_translation_invariance_free_cluster(oracle):
    sysd = lat.cluster(3, 2, 2, x_h=(0.0, 0.3), seed=5)
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    f = oracle.fact_residual(oracle.system(sysd, nnb, nbr))
    total = f["Rq"].sum(0)
    assert np.abs(total).max() < 1e-10 * np.abs(f["Rq"]).max() * sysd["n"]


def test_perfect_hcp_symmetry(oracle):
    """The Stroud points lie on the Cartesian axes, so of the hcp site symmetry
    (-6m2) the rule keeps the mirrors x -> -x and z -> -z: on a perfect crystal
    the x and z force components vanish, the y component is the same for every
    site up to sign, and every width residual is equal."""
    sysd = lat.hcp_mgh(5, 3, 3, with_h=False, pos_noise=0.0, sigma_jitter=0.0, seed=0)
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    f = oracle.fact_residual(oracle.system(sysd, nnb, nbr))
    scale = np.abs(f["Rq"]).max()
    assert np.abs(f["Rq"][:, [0, 2]]).max() < 1e-9 * max(scale, 1.0)
    assert np.ptp(np.abs(f["Rq"][:, 1])) < 1e-9 * max(scale, 1.0)
    assert abs(f["Rq"][:, 1].sum()) < 1e-9 * max(scale, 1.0) * sysd["n"]
    assert np.ptp(f["Rs"]) < 1e-9 * np.abs(f["Rs"]).max()


def test_zero_width_limit_is_zero_temperature_density(oracle):
    """sigma -> 0: the mean-field density tends to the zero-T density
    (src/ADP/density.cpp:44-104) at O(sigma^2)."""
    sysd = lat.read_dump(os.path.join(GOLDEN, "Mg-hcp-cube-x5-x5-x5.dump"))
    st, nnb, nbr = oracle.neighbors(sysd, RC)
    errs = []
    for sgv in (1e-2, 1e-3):
        sysd["sigma"][:] = sgv
        s = oracle.system(sysd, nnb, nbr)
        f = oracle.fact_residual(s)
        errs.append(rel(f["mf_rho"], oracle.ref_rho0(s)))
    assert errs[1] < 2e-2 * errs[0] + 1e-13


def test_radial_self_product_and_gradient():
    """The identity the kernels use at pair-rule points and for the j1 = j2
    self terms: for psi(p) = (u(r) p, w(r) p p^T), r = |p|, and the dipole +
    quadrupole form <X,Y> = 1/2 a.a' + 1/2 B:B' - 1/6 trB trB'
    (src/ADP/MgHx-ADP.cpp:1233-1281, 1781-1831),
      <psi, psi> = 1/2 u^2 r^2 + 1/3 w^2 r^4,
    and the gradient of L . psi(p) at fixed L = M psi(p) -- half the gradient
    of <psi, psi> -- is h p with h = 1/2 u (u' r + u) + w r^2 (1/3 w' r + 2/3 w).
    Checked against the bilinear form itself and central differences, with
    polynomial u, w (a sign or factor slip in h fails it)."""
    rng = np.random.default_rng(5)
    cu, cw = rng.normal(size=4), rng.normal(size=4)
    u = lambda r: np.polyval(cu, r)
    w = lambda r: np.polyval(cw, r)
    du = lambda r: np.polyval(np.polyder(cu), r)
    dw = lambda r: np.polyval(np.polyder(cw), r)

    def form(p):
        r = np.linalg.norm(p)
        a, B = u(r) * p, w(r) * np.outer(p, p)
        return 0.5 * a @ a + 0.5 * np.sum(B * B) - np.trace(B) ** 2 / 6.0

    for _ in range(20):
        p = rng.normal(size=3) * 2.0
        r = np.linalg.norm(p)
        assert form(p) == pytest.approx(0.5 * u(r) ** 2 * r ** 2 + w(r) ** 2 * r ** 4 / 3.0,
                                        rel=1e-12)
        h = 0.5 * u(r) * (du(r) * r + u(r)) + w(r) * r ** 2 * (dw(r) * r / 3.0 + 2.0 * w(r) / 3.0)
        e = 1e-5
        g = np.array([(form(p + e * d) - form(p - e * d)) / (4 * e) for d in np.eye(3)])
        assert np.allclose(g, h * p, rtol=1e-7, atol=1e-9 * np.abs(h * p).max())


def test_isotropic_hessian_terms_of_the_diagonal_blocks():
    """The closed forms the diagonal-block kernel uses, for the same psi and
    form as above with its 9 x 3 Jacobian J = dpsi/dp written out (rows: the
    dipole, then the quadrupole entries xx yy zz xy xz yz):
      J^T M J = gam p p^T + alp I, alp = u^2/2 + w^2 r^2,
        gam = u u'/r + u'^2/2 + (w'^2 r^2 + 4 w w' r + w^2)/3;
      sum_a (M psi)_a d2psi_a/dp2 = ang_a p p^T + ang_b I,
        ang_a = u (u'' + u'/r)/2 + w (w + w' r + w'' r^2/3),
        ang_b = u u' r/2 + w r^2 (w' r - w)/3;
      J1^T M J0 + J0^T M J1 = 2 a I + 2 b1 p1 p1^T + 2 b0 p0 p0^T + d (p1 p0^T + p0 p1^T)
        with c = p1.p0, U = u'/r, W = w'/r, t = W r^2 + 2 w,
        a = u1 u0/2 + w1 w0 c, b1 = U1 u0/2 + W1 w0 c, b0 = u1 U0/2 + w1 W0 c,
        d = U1 U0 c/2 + W1 W0 c^2/2 - t1 t0/6 + w1 w0;
    and a weighted Jacobian sum sum_k n_k J(p_k) is fixed by sum n u,
    sum n U p p^T, sum n W p p p and sum n w p.  Checked against the Jacobians
    themselves and central differences of J^T M psi at fixed M psi."""
    rng = np.random.default_rng(8)
    cu, cw = rng.normal(size=4), rng.normal(size=4)
    f = lambda c, k, r: np.polyval(np.polyder(c, k) if k else c, r)
    qi, qj = [0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]

    def psi(p):
        r = np.linalg.norm(p)
        return np.concatenate([f(cu, 0, r) * p, [f(cw, 0, r) * p[i] * p[j] for i, j in zip(qi, qj)]])

    def jac(p):
        r = np.linalg.norm(p)
        u, U, w, W = f(cu, 0, r), f(cu, 1, r) / r, f(cw, 0, r), f(cw, 1, r) / r
        J = np.zeros((9, 3))
        J[:3] = u * np.eye(3) + U * np.outer(p, p)
        for m, (i, j) in enumerate(zip(qi, qj)):
            for g in range(3):
                J[3 + m, g] = W * p[g] * p[i] * p[j] + w * ((i == g) * p[j] + (j == g) * p[i])
        return J

    def M(X):  # the form's matrix applied to the rows of X (9 x k or 9)
        Y = np.array(X, dtype=float)
        Y[:3] *= 0.5
        Y[3:6] = 0.5 * X[3:6] - (X[3] + X[4] + X[5]) / 6.0
        return Y

    I3 = np.eye(3)
    for _ in range(10):
        p, p0 = rng.normal(size=3) * 2.0, rng.normal(size=3) * 2.0
        r, r0 = np.linalg.norm(p), np.linalg.norm(p0)
        u, du, d2u = (f(cu, k, r) for k in range(3))
        w, dw, d2w = (f(cw, k, r) for k in range(3))
        J = jac(p)
        alp = 0.5 * u * u + w * w * r * r
        gam = u * du / r + 0.5 * du * du + (dw * dw * r * r + 4 * w * dw * r + w * w) / 3.0
        ref = J.T @ M(J)
        assert np.allclose(gam * np.outer(p, p) + alp * I3, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        L = M(psi(p))
        e = 1e-5
        H = np.array([(jac(p + e * d).T @ L - jac(p - e * d).T @ L) / (2 * e) for d in I3])
        anga = 0.5 * u * (d2u + du / r) + w * (w + dw * r + d2w * r * r / 3.0)
        angb = 0.5 * u * du * r + w * r * r * (dw * r - w) / 3.0
        assert np.allclose(anga * np.outer(p, p) + angb * I3, H, rtol=0, atol=1e-7 * np.abs(H).max())
        u0, U0, w0, W0 = f(cu, 0, r0), f(cu, 1, r0) / r0, f(cw, 0, r0), f(cw, 1, r0) / r0
        U1, W1 = du / r, dw / r
        c = p @ p0
        t1, t0 = W1 * r * r + 2 * w, W0 * r0 * r0 + 2 * w0
        a = 0.5 * u * u0 + w * w0 * c
        b1, b0 = 0.5 * U1 * u0 + W1 * w0 * c, 0.5 * u * U0 + w * W0 * c
        dd = 0.5 * U1 * U0 * c + 0.5 * W1 * W0 * c * c - t1 * t0 / 6.0 + w * w0
        J0 = jac(p0)
        ref = J.T @ M(J0) + J0.T @ M(J)
        got = 2 * a * I3 + 2 * b1 * np.outer(p, p) + 2 * b0 * np.outer(p0, p0) + \
            dd * (np.outer(p, p0) + np.outer(p0, p))
        assert np.allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # a weighted sum of Jacobians from its 20 values
    ps, ns = rng.normal(size=(7, 3)) * 2.0, rng.random(7)
    S = sum(n * jac(p) for n, p in zip(ns, ps))
    z_u = sum(n * f(cu, 0, np.linalg.norm(p)) for n, p in zip(ns, ps))
    z_U = sum(n * f(cu, 1, np.linalg.norm(p)) / np.linalg.norm(p) * np.outer(p, p) for n, p in zip(ns, ps))
    z_W = sum(n * f(cw, 1, np.linalg.norm(p)) / np.linalg.norm(p) * np.einsum("i,j,k->ijk", p, p, p)
              for n, p in zip(ns, ps))
    z_w = sum(n * f(cw, 0, np.linalg.norm(p)) * p for n, p in zip(ns, ps))
    T = np.zeros((9, 3))
    T[:3] = z_u * I3 + z_U
    for m, (i, j) in enumerate(zip(qi, qj)):
        for g in range(3):
            T[3 + m, g] = z_W[i, j, g] + (i == g) * z_w[j] + (j == g) * z_w[i]
    assert np.allclose(T, S, rtol=0, atol=1e-12 * np.abs(S).max())
