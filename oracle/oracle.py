"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module; the product path (``dmdgpu``) never does.

The C sources restate the reference's mean-field DMD path in two forms
(reference order and factorised); see ``dmd_oracle.h`` for the files each
function follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    """Compile liboracle.so in place (gcc; see oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class _System(C.Structure):
    _fields_ = [("n", C.c_int), ("q", _dp), ("sig", _dp), ("x", _dp), ("spc", _ip),
                ("beta", _dp), ("gamma", _dp), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("periodic", C.c_int * 3), ("cap", C.c_int), ("nnb", _ip), ("nbr", _ip),
                ("rule", C.c_int)]


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Oracle:
    """FP64 CPU restatement bound to one set of ADP tables."""

    MAXNEIGH = 250

    def __init__(self, table_dir):
        if not os.path.exists(LIB):
            build()
        self.lib = C.CDLL(LIB)
        L = self.lib
        L.or_load_tables.argtypes = [C.POINTER(C.c_void_p), C.c_char_p]
        L.or_free_tables.argtypes = [C.c_void_p]
        L.or_spline_eval.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double]
        L.or_spline_eval.restype = C.c_double
        L.or_spline_size.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.or_neighbors.argtypes = [C.c_int, _dp, _dp, _dp, _ip, C.c_double, C.c_int, _ip, _ip]
        L.or_diff_neighbors.argtypes = [C.c_int, _dp, _ip, _dp, _dp, _ip, C.c_double, C.c_int,
                                        _ip, _ip]
        L.or_diff_mu.argtypes = [C.c_int, _dp, _ip, _dp, _dp, _dp]
        L.or_diff_rates.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip,
                                    C.c_int, C.c_double, C.c_double, _dp]
        sp = C.POINTER(_System)
        for name in ("or_ref_rho0_i", "or_ref_mf_rho_i"):
            getattr(L, name).argtypes = [sp, C.c_void_p, C.c_int]
            getattr(L, name).restype = C.c_double
        for name in ("or_ref_V0_i", "or_ref_S0_i"):
            getattr(L, name).argtypes = [sp, C.c_void_p, C.c_int, _dp]
            getattr(L, name).restype = C.c_double
        L.or_ref_DV0_i_Dq_u.argtypes = [sp, C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.or_ref_DV0_i_Dxi_u.argtypes = [sp, C.c_void_p, C.c_int, C.c_int, _dp]
        L.or_ref_DV0_i_Dxi_u.restype = C.c_double
        L.or_ref_residual.argtypes = [sp, C.c_void_p, C.c_int, _ip, _dp, _dp, _dp, _dp]
        L.or_fact_phase1.argtypes = [sp, C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.or_fact_residual.argtypes = [sp, C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_ref_D2V0_i_Dq2_u.argtypes = [sp, C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.or_ref_diag_blocks.argtypes = [sp, C.c_void_p, C.c_int, _ip, _dp, _dp]
        L.or_fact_diag_blocks.argtypes = [sp, C.c_void_p, C.c_int, C.c_int, _dp]
        L.or_fn_eval.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, _dp]
        L.or_ref_DV_i_Dq_u.argtypes = [sp, C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.or_ref_D2V_i_Dq2_u.argtypes = [sp, C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.or_ref_zero_t.argtypes = [sp, C.c_void_p, C.c_int, _ip, _dp, _dp, _dp]
        h = C.c_void_p()
        rc = L.or_load_tables(C.byref(h), table_dir.encode())
        if rc != 0:
            raise RuntimeError(f"oracle: cannot load ADP tables from {table_dir}")
        self.tab = h

    def __del__(self):
        try:
            self.lib.or_free_tables(self.tab)
        except Exception:
            pass

    # -- tables -------------------------------------------------------------
    def spline(self, which, slot, deriv, x):
        return self.lib.or_spline_eval(self.tab, which, slot, deriv, float(x))

    def spline_size(self, which, slot):
        dx = C.c_double()
        n = self.lib.or_spline_size(self.tab, which, slot, C.byref(dx))
        return n, dx.value

    # -- neighbours ---------------------------------------------------------
    def neighbors(self, sysd, rc):
        n = sysd["n"]
        q = np.ascontiguousarray(sysd["q"], dtype=np.float64)
        lo = np.ascontiguousarray(sysd["lo"], dtype=np.float64)
        hi = np.ascontiguousarray(sysd["hi"], dtype=np.float64)
        per = np.ascontiguousarray(sysd["periodic"], dtype=np.int32)
        nnb = np.zeros(n, dtype=np.int32)
        nbr = np.zeros((n, self.MAXNEIGH), dtype=np.int32)
        st = self.lib.or_neighbors(n, _d(q), _d(lo), _d(hi), _i(per), rc, self.MAXNEIGH,
                                   _i(nnb), _i(nbr))
        return st, nnb, nbr

    def system(self, sysd, nnb, nbr, rule=3):
        """Pack a system dict (see dmdgpu.lattice) into the C struct.  The
        returned object keeps the arrays alive.  rule: 3 (the reference's
        12-point pair rule), 5 (27-point tensor Gauss-Hermite), 44 (the
        reference's 44-point 6D rule) or 6 (6-point Stroud rule on the
        relative coordinate) for the factorised functions."""
        keep = {}
        for k in ("q", "sigma", "x", "beta", "gamma"):
            keep[k] = np.ascontiguousarray(sysd[k], dtype=np.float64)
        keep["spc"] = np.ascontiguousarray(sysd["species"], dtype=np.int32)
        keep["nnb"] = np.ascontiguousarray(nnb, dtype=np.int32)
        keep["nbr"] = np.ascontiguousarray(nbr, dtype=np.int32)
        s = _System()
        s.n = int(sysd["n"])
        s.q, s.sig, s.x = _d(keep["q"]), _d(keep["sigma"]), _d(keep["x"])
        s.spc = _i(keep["spc"])
        s.beta, s.gamma = _d(keep["beta"]), _d(keep["gamma"])
        for a in range(3):
            s.lo[a] = float(sysd["lo"][a])
            s.hi[a] = float(sysd["hi"][a])
            s.periodic[a] = int(sysd["periodic"][a])
        s.cap = keep["nbr"].shape[1]
        s.nnb, s.nbr = _i(keep["nnb"]), _i(keep["nbr"])
        s.rule = int(rule)
        s._keep = keep
        return s

    # -- reference order ----------------------------------------------------
    def ref_mf_rho(self, s):
        return np.array([self.lib.or_ref_mf_rho_i(C.byref(s), self.tab, i) for i in range(s.n)])

    def ref_rho0(self, s):
        return np.array([self.lib.or_ref_rho0_i(C.byref(s), self.tab, i) for i in range(s.n)])

    def ref_V0(self, s, mf_rho, sites=None):
        mf_rho = np.ascontiguousarray(mf_rho, dtype=np.float64)
        sites = range(s.n) if sites is None else sites
        return np.array([self.lib.or_ref_V0_i(C.byref(s), self.tab, i, _d(mf_rho)) for i in sites])

    def ref_S0(self, s, mf_rho, sites=None):
        mf_rho = np.ascontiguousarray(mf_rho, dtype=np.float64)
        sites = range(s.n) if sites is None else sites
        return np.array([self.lib.or_ref_S0_i(C.byref(s), self.tab, i, _d(mf_rho)) for i in sites])

    def ref_DV0_Dq(self, s, u, i, mf_rho):
        mf_rho = np.ascontiguousarray(mf_rho, dtype=np.float64)
        out = np.zeros(4)
        self.lib.or_ref_DV0_i_Dq_u(C.byref(s), self.tab, u, i, _d(mf_rho), _d(out))
        return out

    def ref_DV0_Dxi(self, s, u, i, mf_rho):
        mf_rho = np.ascontiguousarray(mf_rho, dtype=np.float64)
        return self.lib.or_ref_DV0_i_Dxi_u(C.byref(s), self.tab, u, i, _d(mf_rho))

    def ref_residual(self, s, mf_rho, sites=None):
        mf_rho = np.ascontiguousarray(mf_rho, dtype=np.float64)
        if sites is None:
            cnt, sp = s.n, None
            sarr = None
        else:
            sarr = np.ascontiguousarray(sites, dtype=np.int32)
            cnt, sp = len(sarr), _i(sarr)
        Rq = np.zeros((cnt, 3))
        Rs = np.zeros(cnt)
        Rx = np.zeros(cnt)
        self.lib.or_ref_residual(C.byref(s), self.tab, cnt, sp, _d(mf_rho), _d(Rq), _d(Rs), _d(Rx))
        return Rq, Rs, Rx

    def ref_D2V0_Dq2(self, s, u, i, mf_rho):
        mf_rho = np.ascontiguousarray(mf_rho, dtype=np.float64)
        out = np.zeros(9)
        self.lib.or_ref_D2V0_i_Dq2_u(C.byref(s), self.tab, u, i, _d(mf_rho), _d(out))
        return out.reshape(3, 3)

    def ref_diag_blocks(self, s, mf_rho, sites=None):
        """D_u = sum_i beta_i d2V0_i/dq_u2 for the listed sites, (cnt, 3, 3)."""
        mf_rho = np.ascontiguousarray(mf_rho, dtype=np.float64)
        if sites is None:
            cnt, sp = s.n, None
        else:
            sarr = np.ascontiguousarray(sites, dtype=np.int32)
            cnt, sp = len(sarr), _i(sarr)
        D = np.zeros((cnt, 9))
        self.lib.or_ref_diag_blocks(C.byref(s), self.tab, cnt, sp, _d(mf_rho), _d(D))
        return D.reshape(cnt, 3, 3)

    def ref_zero_t(self, s, rho0, sites=None):
        """Zero-temperature force and diagonal blocks with the residual's
        weights: (F (cnt, 3), D (cnt, 3, 3)); rho0 = ref_rho0(s)."""
        rho0 = np.ascontiguousarray(rho0, dtype=np.float64)
        if sites is None:
            cnt, sp = s.n, None
        else:
            sarr = np.ascontiguousarray(sites, dtype=np.int32)
            cnt, sp = len(sarr), _i(sarr)
        F = np.zeros((cnt, 3))
        D = np.zeros((cnt, 9))
        self.lib.or_ref_zero_t(C.byref(s), self.tab, cnt, sp, _d(rho0), _d(F), _d(D))
        return F, D.reshape(cnt, 3, 3)

    def ref_DV_Dq_zero_t(self, s, u, i, rho0):
        rho0 = np.ascontiguousarray(rho0, dtype=np.float64)
        out = np.zeros(3)
        self.lib.or_ref_DV_i_Dq_u(C.byref(s), self.tab, u, i, _d(rho0), _d(out))
        return out

    def ref_D2V_Dq2_zero_t(self, s, u, i, rho0):
        rho0 = np.ascontiguousarray(rho0, dtype=np.float64)
        out = np.zeros(9)
        self.lib.or_ref_D2V_i_Dq2_u(C.byref(s), self.tab, u, i, _d(rho0), _d(out))
        return out.reshape(3, 3)

    def fn_eval(self, which, deriv, code, n, q, slots):
        """One ADP pair/triplet function (which: 0 rho, 1 pair, 2 dipole,
        3 quadrupole) at explicit positions q (ns x 3)."""
        n = np.ascontiguousarray(n, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64).ravel()
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        out = np.zeros(9)
        self.lib.or_fn_eval(self.tab, which, deriv, code, _d(n), _d(q), _i(sl), _d(out))
        return out[0] if deriv == 0 else (out[:3] if deriv == 1 else out.reshape(3, 3))

    # -- factorised ----------------------------------------------------------
    def fact_diag_blocks(self, s, exact=False, nthreads=0):
        D = np.zeros((s.n, 9))
        self.lib.or_fact_diag_blocks(C.byref(s), self.tab, nthreads, int(bool(exact)), _d(D))
        return D.reshape(s.n, 3, 3)

    def fact_residual(self, s, nthreads=0):
        n = s.n
        out = {k: np.zeros(n) for k in ("Rs", "Rx", "S0", "V0", "mf_rho")}
        out["Rq"] = np.zeros((n, 3))
        tot = np.zeros(2)
        self.lib.or_fact_residual(C.byref(s), self.tab, nthreads, _d(out["Rq"]), _d(out["Rs"]),
                                  _d(out["Rx"]), _d(out["S0"]), _d(out["V0"]), _d(out["mf_rho"]),
                                  _d(tot))
        out["S0_total"], out["V0_total"] = tot[0], tot[1]
        return out

    # -- diffusion step (build-defined rate law) ---------------------------------
    DCAP = 50       # max_chemical_neighs, src/Macros.hpp:863
    RC_EB = 2.6     # r_cutoff_Eb, src/Macros.hpp:860
    A2_PER_M2 = 1e20

    def diff_neighbors(self, sysd, rc=None):
        n = sysd["n"]
        q = np.ascontiguousarray(sysd["q"], dtype=np.float64)
        spc = np.ascontiguousarray(sysd["species"], dtype=np.int32)
        lo = np.ascontiguousarray(sysd["lo"], dtype=np.float64)
        hi = np.ascontiguousarray(sysd["hi"], dtype=np.float64)
        per = np.ascontiguousarray(sysd["periodic"], dtype=np.int32)
        cnt = np.zeros(n, dtype=np.int32)
        lst = np.full((n, self.DCAP), -1, dtype=np.int32)
        st = self.lib.or_diff_neighbors(n, _d(q), _i(spc), _d(lo), _d(hi), _i(per),
                                        self.RC_EB if rc is None else rc, self.DCAP, _i(cnt),
                                        _i(lst))
        return st, cnt, lst

    def diff_mu(self, sysd, Rx):
        n = sysd["n"]
        x = np.ascontiguousarray(sysd["x"], dtype=np.float64)
        spc = np.ascontiguousarray(sysd["species"], dtype=np.int32)
        gam = np.ascontiguousarray(sysd["gamma"], dtype=np.float64)
        Rx = np.ascontiguousarray(Rx, dtype=np.float64)
        mu = np.zeros(n)
        self.lib.or_diff_mu(n, _d(x), _i(spc), _d(gam), _d(Rx), _d(mu))
        return mu

    def diff_rates(self, sysd, mu, cnt, lst, D0=1e-7, Eb=0.4, xa=None):
        """xdot per site (1/s) for D0 in m^2/s and Eb in eV; xa (default x)
        decides which sites count as active for the mobility."""
        n = sysd["n"]
        arr = {k: np.ascontiguousarray(sysd[k], dtype=np.float64)
               for k in ("q", "x", "beta", "lo", "hi")}
        xa = arr["x"] if xa is None else np.ascontiguousarray(xa, dtype=np.float64)
        per = np.ascontiguousarray(sysd["periodic"], dtype=np.int32)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        cnt = np.ascontiguousarray(cnt, dtype=np.int32)
        lst = np.ascontiguousarray(lst, dtype=np.int32)
        xd = np.zeros(n)
        self.lib.or_diff_rates(n, _d(arr["q"]), _d(arr["x"]), _d(xa), _d(mu), _d(arr["beta"]),
                               _d(arr["lo"]), _d(arr["hi"]), _i(per), _i(cnt), _i(lst),
                               lst.shape[1], D0 * self.A2_PER_M2, Eb, _d(xd))
        return xd

    def diffusion_step(self, sysd, nnb, nbr, dt=1e-9, D0=1e-7, Eb=0.4, tol=1e-10, max_iter=60,
                       nthreads=0):
        """Backward-Euler occupancy step x = x_n + dt xdot(x) by Picard sweeps
        (positions and widths fixed; which sites are active is frozen at x_n),
        as the GPU solve does.  Returns (x, sweeps, last max |dx|)."""
        st, cnt, lst = self.diff_neighbors(sysd)
        if st:
            raise ValueError(f"diffusive neighbour lists: status {st}")
        h = sysd["species"] == 1
        xn = np.asarray(sysd["x"], dtype=np.float64).copy()
        act = xn >= 5e-4
        x = xn.copy()
        for it in range(1, max_iter + 1):
            s = dict(sysd)
            s["x"] = np.where(act, x, 0.0)
            if (s["x"][act] < 5e-4).any():
                raise ValueError("an active site fell below 5e-4 within the step")
            res = self.fact_residual(self.system(s, nnb, nbr), nthreads)
            mu = self.diff_mu(s, res["Rx"])
            s["x"] = x
            xd = self.diff_rates(s, mu, cnt, lst, D0, Eb, xa=xn)
            xt = np.where(h, xn + dt * xd, x)
            dmax = float(np.abs(xt - x).max())
            x = xt
            if dmax <= tol:
                break
        return x, it, dmax
