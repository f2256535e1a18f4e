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


# ---------------------------------------------------------------------------
# setfl (LAMMPS pair_style adp) -> reference tables.  A re-derivation of the
# reference's converter external/nist2solera/nist2solera.cpp: natural cubic
# splines through the tabulated points (:53-103), r*phi -> phi with a cubic
# extrapolation to r = 0 (:214-229), a zero appended at r_c when the grid
# stops short of it by 1e-4 or more (:151-159), and its fixed-point writer
# (:311-435).  Files are named as the reference's loader opens them
# (adpmg.dat, adph.dat, adpmgh.dat: lower case).
# ---------------------------------------------------------------------------

def read_setfl_adp(path):
    """Parse an ADP setfl file: 3 comment lines; N and the element names;
    ``Nrho drho Nr dr rc``; per element ``Z mass a lattice``, F(rho) (Nrho
    values), rho(r) (Nr values); then r*phi(r), u(r) and w(r) for the pairs
    i <= j in j-major order (nist2solera.cpp:120-211)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    tok = " ".join(lines[3:]).split()
    pos = 0

    def take(k):
        nonlocal pos
        out = tok[pos:pos + k]
        pos += k
        return out

    n_el = int(take(1)[0])
    names = take(n_el)
    n_rho, d_rho, n_r, d_r, rc = take(5)
    n_rho, n_r, d_rho, d_r, rc = int(n_rho), int(n_r), float(d_rho), float(d_r), float(rc)
    el = []
    for _ in range(n_el):
        z, mass, a, lat = take(4)
        F = np.array(take(n_rho), dtype=np.float64)
        rho = np.array(take(n_r), dtype=np.float64)
        el.append({"Z": int(z), "mass": float(mass), "a": float(a), "lattice": lat, "F": F,
                   "rho": rho})
    pairs = [(i, j) for j in range(n_el) for i in range(j + 1)]
    tabs = {}
    for key in ("rphi", "u", "w"):
        for ij in pairs:
            tabs[(key,) + ij] = np.array(take(n_r), dtype=np.float64)
    return {"names": names, "n_rho": n_rho, "d_rho": d_rho, "n_r": n_r, "d_r": d_r, "rc": rc,
            "elements": el, "pairs": pairs, "tables": tabs}


def write_setfl_adp(path, names, masses, n_rho, d_rho, n_r, d_r, rc, F, rho, rphi, u, w):
    """Write an ADP setfl file in the layout read_setfl_adp parses (synthetic
    potentials for tests and examples).  F[i], rho[i]: per element; rphi, u, w:
    dicts keyed by (i, j), i <= j."""
    out = ["synthetic ADP potential", "written by dmdgpu.tables", "setfl layout",
           f" {len(names)} " + " ".join(names),
           f"{n_rho} {d_rho!r} {n_r} {d_r!r} {rc!r}"]
    for i in range(len(names)):
        out.append(f"{i + 1} {masses[i]!r} 1.0 hcp")
        out += [repr(float(v)) for v in F[i]]
        out += [repr(float(v)) for v in rho[i]]
    for tab in (rphi, u, w):
        for j in range(len(names)):
            for i in range(j + 1):
                out += [repr(float(v)) for v in tab[(i, j)]]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def natural_cubic_spline(x, y):
    """Coefficients (a, b, c, d) per segment of the natural cubic spline
    through (x_k, y_k), by the tridiagonal sweep of nist2solera.cpp:53-103
    (the same operations in the same order, so the same doubles)."""
    n = len(x) - 1
    a = [float(v) for v in y]
    x = [float(v) for v in x]
    h = [x[i + 1] - x[i] for i in range(n)]
    alpha = [0.0] * n
    for i in range(1, n):
        alpha[i] = 3 * (a[i + 1] - a[i]) / h[i] - 3 * (a[i] - a[i - 1]) / h[i - 1]
    c = [0.0] * (n + 1)
    l = [0.0] * (n + 1)
    mu = [0.0] * (n + 1)
    z = [0.0] * (n + 1)
    l[0] = 1.0
    for i in range(1, n):
        l[i] = 2 * (x[i + 1] - x[i - 1]) - h[i - 1] * mu[i - 1]
        mu[i] = h[i] / l[i]
        z[i] = (alpha[i] - h[i - 1] * z[i - 1]) / l[i]
    b = [0.0] * n
    d = [0.0] * n
    for j in range(n - 1, -1, -1):
        c[j] = z[j] - mu[j] * c[j + 1]
        b[j] = (a[j + 1] - a[j]) / h[j] - h[j] * (c[j + 1] + 2 * c[j]) / 3
        d[j] = (c[j + 1] - c[j]) / 3 / h[j]
    return a[:n], b, c[:n], d


def _spline_block(x, y, dx):
    a, b, c, d = natural_cubic_spline(x, y)
    out = [f"{len(a)} {dx:.16f}\n"]
    for k in range(len(a)):
        vals = (a[k], b[k], c[k], d[k], b[k], 2 * c[k], 3 * d[k], 2 * c[k], 6 * d[k])
        out.append(" ".join(f"{v:.16f}" for v in vals) + "\n")
    return "".join(out)


def setfl_to_tables(path, folder, tol=1e-4):
    """Convert an ADP setfl file into the reference's tables in ``folder``:
    one ``adp<el>.dat`` per element (rho, F, phi, u, w of the like pair) and
    one ``adp<el_i><el_j>.dat`` per unlike pair (phi, u, w).  Returns the
    written paths."""
    S = read_setfl_adp(path)
    n_r, d_r, rc = S["n_r"], S["d_r"], S["rc"]
    extend = not (rc - (n_r - 1) * d_r < tol)
    cols = n_r + 1 if extend else n_r
    r = [k * d_r for k in range(cols)]
    rho_x = [k * S["d_rho"] for k in range(S["n_rho"])]

    def padded(v):
        v = [float(t) for t in v]
        return v + [0.0] if extend else v

    def phi_of(rphi):
        v = padded(rphi)
        phi = [0.0] * cols
        for k in range(1, cols):
            phi[k] = v[k] / r[k]
        phi[0] = 4 * phi[1] - 6 * phi[2] + 4 * phi[3] - phi[4]
        return phi

    os.makedirs(folder, exist_ok=True)
    names = [nm.lower() for nm in S["names"]]
    written = []
    for (i, j) in S["pairs"]:
        T = S["tables"]
        body = [_spline_block(r, phi_of(T[("rphi", i, j)]), d_r),
                _spline_block(r, padded(T[("u", i, j)]), d_r),
                _spline_block(r, padded(T[("w", i, j)]), d_r)]
        if i == j:
            e = S["elements"][i]
            head = f"{e['mass']:.5f} {1.0:.1f} {rc:.16f} \n"
            body = [_spline_block(r, padded(e["rho"]), d_r),
                    _spline_block(rho_x, e["F"], S["d_rho"])] + body
            fname = f"adp{names[i]}.dat"
        else:
            head = f"{rc:.16f}\n"
            fname = f"adp{names[i]}{names[j]}.dat"
        dst = os.path.join(folder, fname)
        with open(dst, "w") as fh:
            fh.write(head + "".join(body))
        written.append(dst)
    return written


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
