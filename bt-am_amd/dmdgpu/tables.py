"""Writers for ADP spline tables in the reference file layout.

Layout (read by init_adp_MgHx, src/ADP/MgHx-ADP.cpp:366-501):
  element file (adpmg.dat, adph.dat): ``mass factor rcut``, then the blocks
  rho, embed, pair, u, w;  alloy file (adpmgh.dat): ``rcut``, then pair, u, w.
  Each block is ``n dx`` followed by n rows of 9 columns
  ``a b c d db dc dd ddc ddd`` with db = b, dc = 2c, dd = 3d, ddc = 2c, ddd = 6d,
  so that S(x) = a + b p + c p^2 + d p^3 on segment m, p = x - m dx.

Segments are cubic Hermite interpolants of a function and its derivative at
the knots, so the tables are C1 and reproduce cubic polynomials exactly.
"""
import os

import numpy as np

MASS = {"Mg": 24.305, "H": 1.008}  # header masses of the reference tables


def hermite_rows(f, df, n, dx):
    """n x 9 coefficient rows of the C1 cubic Hermite spline of f on [0, n dx]."""
    x0 = np.arange(n) * dx
    x1 = x0 + dx
    f0, f1 = f(x0), f(x1)
    d0, d1 = df(x0), df(x1)
    a = f0
    b = d0
    c = (3.0 * (f1 - f0) / dx - 2.0 * d0 - d1) / dx
    d = (d0 + d1 - 2.0 * (f1 - f0) / dx) / dx ** 2
    return np.stack([a, b, c, d, b, 2 * c, 3 * d, 2 * c, 6 * d], axis=1)


def _block(fh, rows, dx):
    fh.write(f"{rows.shape[0]} {dx:.16f}\n")
    for r in rows:
        fh.write(" ".join(f"{v:.17e}" for v in r) + "\n")


def write_tables(folder, fn, nr, dxr, ne, dxe, rcut=None):
    """Write adpmg.dat / adph.dat / adpmgh.dat into ``folder``.

    ``fn`` maps names to (f, df) pairs: rho_Mg, rho_H, F_Mg, F_H, phi_MgMg,
    phi_HH, phi_MgH, u_MgMg, u_HH, u_MgH, w_MgMg, w_HH, w_MgH."""
    os.makedirs(folder, exist_ok=True)
    rc = nr * dxr if rcut is None else rcut
    for sp, fname in (("Mg", "adpmg.dat"), ("H", "adph.dat")):
        with open(os.path.join(folder, fname), "w") as fh:
            fh.write(f"{MASS[sp]:.5f} 1.0 {rc:.16f}\n")
            _block(fh, hermite_rows(*fn["rho_" + sp], nr, dxr), dxr)
            _block(fh, hermite_rows(*fn["F_" + sp], ne, dxe), dxe)
            pp = sp + sp
            for k in ("phi_", "u_", "w_"):
                _block(fh, hermite_rows(*fn[k + pp], nr, dxr), dxr)
    with open(os.path.join(folder, "adpmgh.dat"), "w") as fh:
        fh.write(f"{rc:.16f}\n")
        for k in ("phi_", "u_", "w_"):
            _block(fh, hermite_rows(*fn[k + "MgH"], nr, dxr), dxr)


def poly(*c):
    """(f, df) of the polynomial c0 + c1 x + c2 x^2 + c3 x^3."""
    c = list(c) + [0.0] * (4 - len(c))
    f = lambda x: c[0] + x * (c[1] + x * (c[2] + x * c[3]))
    df = lambda x: c[1] + x * (2 * c[2] + x * 3 * c[3])
    return f, df


def smooth(amp, decay, r0, rc):
    """(f, df) of amp * exp(-decay (r - r0)) * ((rc - r)/rc)^4 for r < rc, else 0:
    value and slope vanish at rc, so the reference clamp past the table end is
    consistent with the derivative."""
    def f(r):
        r = np.asarray(r, dtype=np.float64)
        t = np.clip((rc - r) / rc, 0.0, None)
        return amp * np.exp(-decay * (r - r0)) * t ** 4

    def df(r):
        r = np.asarray(r, dtype=np.float64)
        t = np.clip((rc - r) / rc, 0.0, None)
        e = amp * np.exp(-decay * (r - r0))
        return e * (-decay * t ** 4 - 4.0 * t ** 3 / rc)
    return f, df


def synthetic_mgh(folder, seed=0, nr=10000, rc=6.5, ne=10000, rho_max=2.0):
    """Seeded smooth Mg-H tables, nr = nrho = 10000, cut-off rc = 6.5 A."""
    rng = np.random.default_rng(seed)
    jit = lambda: 1.0 + 0.1 * (2 * rng.random() - 1)
    fn = {}
    fn["rho_Mg"] = smooth(0.05 * jit(), 1.0 * jit(), 3.2, rc)
    fn["rho_H"] = smooth(0.08 * jit(), 1.4 * jit(), 2.0, rc)
    for sp, a in (("Mg", 1.2), ("H", 0.9)):
        a = a * jit()
        fn["F_" + sp] = (lambda x, a=a: -a * np.sqrt(x + 0.01) + 0.3 * a * x,
                         lambda x, a=a: -0.5 * a / np.sqrt(x + 0.01) + 0.3 * a)
    for pp, (amp, r0) in (("MgMg", (0.25, 3.2)), ("HH", (0.15, 1.5)), ("MgH", (0.2, 2.0))):
        fn["phi_" + pp] = smooth(amp * jit(), 1.6 * jit(), r0, rc)
        fn["u_" + pp] = smooth(0.02 * jit(), 1.1 * jit(), r0, rc)
        fn["w_" + pp] = smooth(0.01 * jit(), 1.2 * jit(), r0, rc)
    write_tables(folder, fn, nr, rc / nr, ne, rho_max / ne)
    return folder
