"""CPU tests of the setfl -> table converter (dmdgpu.tables.setfl_to_tables),
a re-derivation of external/nist2solera/nist2solera.cpp.

Pinned by: the reference's own converted tables (inputs/adpmg.dat, adph.dat,
adpmgh.dat, shipped in bt-am_amd/potentials) from its own setfl input
(inputs/Mg_H.adp.alloy.txt, in tests/golden); scipy's natural cubic spline;
and closed forms of the r = 0 extrapolation and of the appended cut-off point.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, TABLE_DIR

import dmdgpu.tables as tabs
from test_oracle import _raw_blocks

SETFL = os.path.join(GOLDEN, "Mg_H.adp.alloy.txt")


@pytest.mark.parametrize("fname,alloy", [("adpmg.dat", False), ("adph.dat", False),
                                         ("adpmgh.dat", True)])
def test_conversion_reproduces_the_reference_tables(tmp_path, fname, alloy):
    """Same header, segment counts and steps; coefficients within 1e-11 of
    each column's scale in the pair block (measured 8.4e-12 at most) and 5e-16
    elsewhere.  The reference binary
    is built with -Ofast (external/nist2solera/Makefile), whose reciprocal
    multiplications move last bits, and the pair block's r*phi / r near r = 0
    is ill-conditioned through the spline sweep."""
    tabs.setfl_to_tables(SETFL, str(tmp_path))
    mine = os.path.join(str(tmp_path), fname)
    ref = os.path.join(TABLE_DIR, fname)
    assert open(mine).readline() == open(ref).readline()
    A, B = _raw_blocks(mine, alloy), _raw_blocks(ref, alloy)
    assert len(A) == len(B)
    pair_block = 0 if alloy else 2
    for k, ((na, dxa, ra), (nb, dxb, rb)) in enumerate(zip(A, B)):
        assert na == nb and dxa == dxb
        err = (np.abs(ra - rb).max(0) / np.abs(rb).max(0).clip(1e-300)).max()
        assert err < (1e-11 if k == pair_block else 5e-16), (k, err)


def test_natural_spline_matches_scipy():
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(3)
    x = np.arange(40) * 0.05
    y = np.sin(3 * x) + 0.1 * rng.standard_normal(40)
    a, b, c, d = (np.array(v) for v in tabs.natural_cubic_spline(x, y))
    cs = CubicSpline(x, y, bc_type="natural")
    ref = cs.c[::-1]  # scipy: highest power first
    for mine, r in zip((a, b, c, d), ref):
        assert np.abs(mine - r).max() <= 1e-12 * np.abs(r).max()


def test_extrapolation_and_cutoff_point(tmp_path):
    """r*phi(r) = r (1 + r + r^2 + r^3): the cubic extrapolation of phi to
    r = 0 (nist2solera.cpp:214-229) gives phi(0) = 1 exactly; a cut-off 0.5 A
    past the grid appends one zero knot (:151-159), so every r-table gains a
    segment ending at 0."""
    n_r, d_r, n_rho, d_rho = 60, 0.1, 50, 0.02
    r = np.arange(n_r) * d_r
    phi = 1 + r + r ** 2 + r ** 3
    F = [np.linspace(0, -1, n_rho)] * 2
    rho = [np.exp(-r)] * 2
    tab = {(i, j): phi * r for j in range(2) for i in range(j + 1)}
    uw = {(i, j): np.exp(-r) for j in range(2) for i in range(j + 1)}
    for rc, extra in (((n_r - 1) * d_r, 0), ((n_r - 1) * d_r + 0.5, 1)):
        path = os.path.join(str(tmp_path), f"t{extra}.adp")
        tabs.write_setfl_adp(path, ["Mg", "H"], [24.305, 1.008], n_rho, d_rho, n_r, d_r, rc,
                             F, rho, tab, uw, uw)
        out = os.path.join(str(tmp_path), f"o{extra}")
        tabs.setfl_to_tables(path, out)
        blocks = _raw_blocks(os.path.join(out, "adpmg.dat"), False)
        nseg = [b[0] for b in blocks]
        assert nseg == [n_r - 1 + extra, n_rho - 1, n_r - 1 + extra, n_r - 1 + extra,
                        n_r - 1 + extra]
        pair = blocks[2][2]
        assert pair[0, 0] == pytest.approx(1.0, abs=1e-12)  # phi(0)
        if extra:
            rows = blocks[0][2]
            end = rows[-1, 0] + rows[-1, 1] * d_r + rows[-1, 2] * d_r ** 2 + rows[-1, 3] * d_r ** 3
            assert abs(end) < 1e-12  # the appended knot: rho(rc) = 0
