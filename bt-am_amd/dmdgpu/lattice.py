"""Seeded Mg-H geometries and the reference dump reader.

A *system* is a dict of DMSwarm-equivalent SoA arrays (the fields registered in
src/Atoms/Atom.cpp:149-203): ``q`` (n x 3 mean positions, Angstrom), ``sigma``
(n widths), ``x`` (n occupancies / molar fractions), ``species`` (AtomicSpecie
codes, H = 1, Mg = 12), ``beta`` (1 / k_B T) and ``gamma`` (chemical
multiplier), plus the box ``lo``/``hi`` and the ``periodic`` flag of each axis.
"""
import numpy as np

K_B = 8.617332478e-5  # eV/K, src/Macros.hpp:849
MG, H = 12, 1

# Orthorhombic hcp cell (a, sqrt(3) a, c) with four atoms (fractional units).
HCP_BASIS = np.array([[0.0, 0.0, 0.0],
                      [0.5, 0.5, 0.0],
                      [0.5, 5.0 / 6.0, 0.5],
                      [0.0, 1.0 / 3.0, 0.5]])

# Starting widths of the generated systems (the reference stores sigma only in
# its dump files).  The harmonic estimate sigma^2 = k_B T / k with
# k ~ 1.8 eV/A^2 for Mg at 300 K gives ~0.12 A; H sits in a softer, lighter
# well and starts wider.
SIGMA0 = {MG: 0.12, H: 0.15}


def _finish(q, spc, x, lo, hi, periodic, T, rng, sigma_jitter):
    n = len(spc)
    sig = np.array([SIGMA0[s] for s in spc], dtype=np.float64)
    if sigma_jitter > 0:
        sig *= 1.0 + sigma_jitter * (2.0 * rng.random(n) - 1.0)
    beta = np.full(n, 1.0 / (K_B * T))
    return {"n": n, "q": np.ascontiguousarray(q, dtype=np.float64), "sigma": sig,
            "x": np.ascontiguousarray(x, dtype=np.float64),
            "species": np.ascontiguousarray(spc, dtype=np.int32), "beta": beta,
            "gamma": np.zeros(n), "lo": np.asarray(lo, dtype=np.float64),
            "hi": np.asarray(hi, dtype=np.float64),
            "periodic": np.asarray(periodic, dtype=np.int32)}


def hcp_mgh(nx, ny, nz, a=3.209, c_over_a=1.624, with_h=True, x_h=(0.0, 0.1), seed=0,
            pos_noise=0.01, sigma_jitter=0.05, T=300.0, periodic=(1, 1, 1)):
    """hcp Mg replicated nx x ny x nz orthorhombic cells (4 atoms each).

    With ``with_h`` every Mg carries two tetrahedral H sites at +/- 3c/8 along c
    (2 per Mg), with occupancy x ~ U[x_h].  Mean positions get seeded
    N(0, pos_noise^2) noise per coordinate."""
    rng = np.random.default_rng(seed)
    c = a * c_over_a
    cell = np.array([a, np.sqrt(3.0) * a, c])
    idx = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"),
                   axis=-1).reshape(-1, 3)
    mg = ((idx[:, None, :] + HCP_BASIS[None, :, :]) * cell).reshape(-1, 3)
    pos = [mg]
    spc = [np.full(len(mg), MG)]
    x = [np.ones(len(mg))]
    if with_h:
        up = mg + np.array([0.0, 0.0, 3.0 * c / 8.0])
        dn = mg - np.array([0.0, 0.0, 3.0 * c / 8.0])
        hpos = np.stack([up, dn], axis=1).reshape(-1, 3)
        pos.append(hpos)
        spc.append(np.full(len(hpos), H))
        x.append(rng.uniform(x_h[0], x_h[1], len(hpos)))
    q = np.concatenate(pos)
    spc = np.concatenate(spc)
    x = np.concatenate(x)
    q = q + rng.normal(0.0, pos_noise, q.shape)
    lo = np.zeros(3)
    hi = cell * np.array([nx, ny, nz])
    # wrap into the box on periodic axes
    for ax in range(3):
        if periodic[ax]:
            q[:, ax] = lo[ax] + np.mod(q[:, ax] - lo[ax], hi[ax] - lo[ax])
    return _finish(q, spc, x, lo, hi, periodic, T, rng, sigma_jitter)


def cluster(nx, ny, nz, **kw):
    """Free (non-periodic) hcp block: every axis open."""
    kw.setdefault("periodic", (0, 0, 0))
    return hcp_mgh(nx, ny, nz, **kw)


def read_dump(path):
    """Reader for the reference dump layout (src/IO/dump-input.cpp:51-215):
    TIMESTEP, NUMBER OF ATOMS, BOX BOUNDS with pp/ff flags, then rows of
    ``type x y z stdvq molarfraction chemicalmultp thermalmultp ...``."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[3].split()[0])
    flags = lines[4].split()[3:6]
    periodic = [1 if fl == "pp" else 0 for fl in flags]
    for fl in flags:
        if fl not in ("pp", "ff"):
            raise ValueError(f"unrecognised boundary flag {fl}")
    lo, hi = [], []
    for k in range(3):
        a, b = lines[5 + k].split()[:2]
        lo.append(float(a))
        hi.append(float(b))
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[9:9 + n]])
    spc = rows[:, 0].astype(np.int32)
    bad = set(spc.tolist()) - {MG, H}
    if bad:
        raise ValueError(f"species {sorted(bad)} not supported (Mg-H only)")
    return {"n": n, "q": np.ascontiguousarray(rows[:, 1:4]), "sigma": rows[:, 4].copy(),
            "x": rows[:, 5].copy(), "species": spc, "beta": rows[:, 7].copy(),
            "gamma": rows[:, 6].copy(), "lo": np.array(lo), "hi": np.array(hi),
            "periodic": np.array(periodic, dtype=np.int32)}


def _hcp_region(keep, nx, ny, nz, a, c_over_a, x_h, seed, pos_noise, sigma_jitter, T, periodic):
    """hcp Mg (+ tetrahedral H) of an nx x ny x nz block, keeping the sites for
    which keep(pos) is true; the box spans the block."""
    full = hcp_mgh(nx, ny, nz, a=a, c_over_a=c_over_a, with_h=True, x_h=x_h, seed=seed,
                   pos_noise=0.0, sigma_jitter=0.0, T=T, periodic=periodic)
    rng = np.random.default_rng(seed + 7919)
    m = keep(full["q"])
    q = full["q"][m] + rng.normal(0.0, pos_noise, (int(m.sum()), 3))
    spc = full["species"][m]
    x = full["x"][m]
    for ax in range(3):
        if periodic[ax]:
            L = full["hi"][ax] - full["lo"][ax]
            q[:, ax] = full["lo"][ax] + np.mod(q[:, ax] - full["lo"][ax], L)
    return _finish(q, spc, x, full["lo"], full["hi"], periodic, T, rng, sigma_jitter)


def nanowire(radius, ncells_c=10, a=3.209, c_over_a=1.624, x_h=(0.0, 0.3), seed=0,
             pos_noise=0.01, sigma_jitter=0.05, T=300.0):
    """Circular hcp Mg-H nanowire along c of the given radius (A), ncells_c
    cells long and periodic along its axis only (free surfaces elsewhere)."""
    c = a * c_over_a
    nx = int(np.ceil(2 * radius / a)) + 2
    ny = int(np.ceil(2 * radius / (np.sqrt(3) * a))) + 2
    cx, cy = nx * a / 2, ny * np.sqrt(3) * a / 2
    keep = lambda q: (q[:, 0] - cx) ** 2 + (q[:, 1] - cy) ** 2 <= radius ** 2
    return _hcp_region(keep, nx, ny, ncells_c, a, c_over_a, x_h, seed, pos_noise, sigma_jitter, T,
                       (0, 0, 1))


def sphere(radius, a=3.209, c_over_a=1.624, x_h=(0.0, 0.3), seed=0, pos_noise=0.01,
           sigma_jitter=0.05, T=300.0):
    """Free hcp Mg-H sphere of the given radius (A)."""
    c = a * c_over_a
    nx = int(np.ceil(2 * radius / a)) + 2
    ny = int(np.ceil(2 * radius / (np.sqrt(3) * a))) + 2
    nz = int(np.ceil(2 * radius / c)) + 2
    ctr = np.array([nx * a, ny * np.sqrt(3) * a, nz * c]) / 2
    keep = lambda q: ((q - ctr) ** 2).sum(1) <= radius ** 2
    return _hcp_region(keep, nx, ny, nz, a, c_over_a, x_h, seed, pos_noise, sigma_jitter, T,
                       (0, 0, 0))


def mixed_batch(count=64, seed=2024):
    """The batched-relaxation workload: count/2 Mg-H nanowires along c
    (R ~ U[15, 30] A, 10 cells long, periodic along the axis) and count/2 free
    spheres (R ~ U[12, 25] A), H occupancy x ~ U[0, 0.3], one seed per
    instance.  Sizes run from about 900 to 19,000 sites."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        s = int(rng.integers(1 << 30))
        if k % 2 == 0:
            out.append(nanowire(float(rng.uniform(15.0, 30.0)), seed=s))
        else:
            out.append(sphere(float(rng.uniform(12.0, 25.0)), seed=s))
    return out


def box_of_size(sites, seed=0, x_h=(0.0, 0.1)):
    """Periodic hcp Mg-H box of about `sites` sites (12 per orthorhombic cell),
    near-cubic, at least 4 x 3 x 3 cells (every edge >= 2 r_c)."""
    a, c = 3.209, 3.209 * 1.624
    side = (sites / 12.0 * a * np.sqrt(3.0) * a * c) ** (1.0 / 3.0)
    nx = max(4, int(round(side / a)))
    ny = max(3, int(round(side / (np.sqrt(3.0) * a))))
    nz = max(3, int(round(side / c)))
    return hcp_mgh(nx, ny, nz, x_h=x_h, seed=seed)


def wire_of_size(sites, seed=0, x_h=(0.0, 0.3)):
    """Nanowire (10 cells along c, periodic along c) of about `sites` sites:
    about 21.1 R^2 sites for radius R in A."""
    return nanowire(float(np.sqrt(sites / 21.1)), seed=seed, x_h=x_h)


def serving_mix(count, seed=7, min_sites=500, max_sites=200_000, diffuse_frac=0.5):
    """Request mix of the serving loop: (kind, system) pairs, kind "relax" or
    "diffuse", sizes log-uniform in [min_sites, max_sites], half periodic
    boxes and half nanowires, one seed per request."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = float(np.exp(rng.uniform(np.log(min_sites), np.log(max_sites))))
        s = int(rng.integers(1 << 30))
        sysd = box_of_size(n, seed=s) if rng.random() < 0.5 else wire_of_size(n, seed=s)
        kind = "diffuse" if rng.random() < diffuse_frac else "relax"
        out.append((kind, sysd))
    return out


def write_dump(path, sysd, step=0):
    """Writes a system in the reference's dump layout (the layout
    src/IO/dump-input.cpp:51-215 reads; writer src/IO/dump-output.cpp:52-159):
    TIMESTEP, NUMBER OF ATOMS, BOX BOUNDS with the pp/ff flag of each axis,
    then ``type x y z stdvq molarfraction chemicalmultp thermalmultp
    chemicalmultpbcc thermalmultpbcc`` per site.  Two differences from the
    reference writer, so that the file reads back: the species code is written
    as is (the writer adds 1, which its reader rejects) and each axis keeps its
    own boundary flag (the writer always writes pp).  Values are written in their shortest
    exact decimal form (the file reads back bit for bit)."""
    n = int(sysd["n"])
    r = lambda v: repr(float(v))  # shortest exact decimal form
    flags = " ".join("pp" if p else "ff" for p in sysd["periodic"])
    lines = ["ITEM: TIMESTEP", str(step), "ITEM: NUMBER OF ATOMS", str(n),
             f"ITEM: BOX BOUNDS {flags}"]
    for a in range(3):
        lines.append(f"{r(sysd['lo'][a])} {r(sysd['hi'][a])}")
    lines.append("ITEM: ATOMS type x y z stdvq molarfraction chemicalmultp thermalmultp "
                 "chemicalmultpbcc thermalmultpbcc")
    q, sg, x = sysd["q"], sysd["sigma"], sysd["x"]
    for i in range(n):
        lines.append(f"{int(sysd['species'][i])} {r(q[i, 0])} {r(q[i, 1])} {r(q[i, 2])} "
                     f"{r(sg[i])} {r(x[i])} {r(sysd['gamma'][i])} {r(sysd['beta'][i])} 0.0 0.0")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
