"""Python host binding of libdmdgpu.so (the C ABI in include/dmdgpu.h).

The operator surface mirrors the reference's per-site evaluators and its SNES
callback (src/ADP/MgHx-mf-V-bulk.cpp, src/Mechanical-eqs/
Mechanical-Relaxation-bulk.cpp:441); every call runs the HIP kernels of
``csrc/dmd_kernels.hip``.  There is no CPU fallback: constructing a
``Context`` without the built library or without a GPU raises.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libdmdgpu.so")
TABLE_DIR = os.path.join(PKG, "potentials")
R_CUTOFF = 6.2934034034  # r_cutoff_ADP_MgHx, src/Macros.hpp:857
MAXNEIGH = 250           # src/Macros.hpp:862
FP64, SPLIT32 = 0, 1     # dmdgpu_set_precision codes

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_IO", 4: "ERR_SPECIES", 5: "ERR_MAXNEIGH",
          6: "ERR_NONEIGH", 7: "ERR_BOX", 8: "ERR_STATE", 9: "ERR_TABLES", 10: "ERR_BUDGET",
          11: "ERR_DOMAIN"}
RELAX_STATE = {0: "idle", 1: "running", 2: "converged", 3: "maxiter", 4: "stalled",
               5: "domain"}

# Every symbol include/dmdgpu.h declares.
EXPORTS = ["dmdgpu_create", "dmdgpu_destroy", "dmdgpu_last_error", "dmdgpu_stream", "dmdgpu_sync",
           "dmdgpu_init_adp_MgHx", "dmdgpu_set_cutoff", "dmdgpu_set_state",
           "dmdgpu_update_state", "dmdgpu_neighbors", "dmdgpu_get_neighbors", "dmdgpu_mf_rho",
           "dmdgpu_site_energies", "dmdgpu_residual", "dmdgpu_residual_resident",
           "dmdgpu_set_tuning", "dmdgpu_set_precision", "dmdgpu_set_timing", "dmdgpu_kernel_count", "dmdgpu_kernel_name",
           "dmdgpu_kernel_times", "dmdgpu_reset_kernel_times", "dmdgpu_work_counts",
           "dmdgpu_add_instance", "dmdgpu_remove_instance", "dmdgpu_instance_info",
           "dmdgpu_set_budget", "dmdgpu_bytes_per_site", "dmdgpu_evaluate",
           "dmdgpu_instance_residual", "dmdgpu_instance_state", "dmdgpu_instance_set_unknowns",
           "dmdgpu_relax_start", "dmdgpu_relax_tick", "dmdgpu_relax_status", "dmdgpu_relax",
           "dmdgpu_set_graph", "dmdgpu_diffuse_start", "dmdgpu_diffusion_rates",
           "dmdgpu_diffusion_neighbors", "dmdgpu_jvp", "dmdgpu_set_rule", "dmdgpu_diag_blocks", "dmdgpu_instance_diag_blocks",
           "dmdgpu_set_tuning_precision", "dmdgpu_device_bytes", "dmdgpu_diag_blocks_noalloc",
           "dmdgpu_site_DV0_Dq", "dmdgpu_site_DV0_Dxi", "dmdgpu_residual_packed",
           "dmdgpu_relax_energy", "dmdgpu_relax_evals", "dmdgpu_instance_pairs",
           "dmdgpu_jvp_packed"]
DCAP = 50  # diffusive neighbours per site, src/Macros.hpp:863

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class RelaxOpts(C.Structure):
    """dmdgpu_relax_opts."""
    _fields_ = [("atol", C.c_double), ("max_newton", C.c_int), ("max_cg", C.c_int),
                ("eta", C.c_double), ("mixed", C.c_int), ("max_evals", C.c_int),
                ("dof_mask", C.c_int), ("probe", C.c_double), ("precond", C.c_int),
                ("merit", C.c_int), ("solver", C.c_int), ("ngmres_m", C.c_int)]


DOF_MG_Q, DOF_MG_SIGMA, DOF_H_Q, DOF_H_SIGMA = 1, 2, 4, 8


def relax_opts(atol=1e-8, max_newton=50, max_cg=60, eta=0.05, mixed=False, max_evals=0,
               dof_mask=DOF_MG_Q | DOF_MG_SIGMA, probe=0.0, precond=False, merit="residual",
               solver="newton", ngmres_m=3):
    """merit: "residual" (line search on |R|_2) or "energy" (on sum_i S0_i).
    solver: "newton" (inexact Newton-CG) or "ngmres" (NGMRES(ngmres_m) with a
    critical-point line search, the reference's SNES configuration)."""
    return RelaxOpts(atol, max_newton, max_cg, eta, 1 if mixed else 0, max_evals, dof_mask,
                     probe, int(precond), {"residual": 0, "energy": 1}[merit],
                     {"newton": 0, "ngmres": 1}[solver], ngmres_m)


class DiffuseOpts(C.Structure):
    """dmdgpu_diffuse_opts."""
    _fields_ = [("dt", C.c_double), ("D0", C.c_double), ("Eb", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int), ("mixed", C.c_int)]


def diffuse_opts(dt=1e-9, D0=1e-7, Eb=0.4, tol=1e-10, max_iter=60, mixed=False):
    """One implicit occupancy step: dt (s), D0 (m^2/s), Eb (eV)."""
    return DiffuseOpts(dt, D0, Eb, tol, max_iter, 1 if mixed else 0)


class DmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def build():
    """Compile libdmdgpu.so for gfx950 in place (make -C bt-am_amd)."""
    subprocess.run(["make", "-s", "-C", PKG], check=True)


_lib = None


def load_library():
    """Load libdmdgpu.so and declare its prototypes (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run dmdgpu.build() first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.dmdgpu_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.dmdgpu_destroy.argtypes = [vp]
    L.dmdgpu_last_error.restype = C.c_char_p
    L.dmdgpu_stream.argtypes = [vp]
    L.dmdgpu_stream.restype = vp
    L.dmdgpu_sync.argtypes = [vp]
    L.dmdgpu_init_adp_MgHx.argtypes = [vp, C.c_int, C.c_char_p]
    L.dmdgpu_set_cutoff.argtypes = [vp, C.c_double]
    L.dmdgpu_set_state.argtypes = [vp, C.c_int, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _ip]
    L.dmdgpu_update_state.argtypes = [vp, _dp, _dp, _dp]
    L.dmdgpu_neighbors.argtypes = [vp]
    L.dmdgpu_get_neighbors.argtypes = [vp, _ip, _ip, C.c_int]
    L.dmdgpu_mf_rho.argtypes = [vp, _dp]
    L.dmdgpu_site_energies.argtypes = [vp, _dp, _dp]
    L.dmdgpu_residual.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.dmdgpu_residual_resident.argtypes = [vp]
    L.dmdgpu_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.dmdgpu_set_precision.argtypes = [vp, C.c_int]
    L.dmdgpu_set_timing.argtypes = [vp, C.c_int]
    L.dmdgpu_kernel_count.restype = C.c_int
    L.dmdgpu_kernel_name.argtypes = [C.c_int]
    L.dmdgpu_kernel_name.restype = C.c_char_p
    L.dmdgpu_kernel_times.argtypes = [vp, _lp, _dp]
    L.dmdgpu_reset_kernel_times.argtypes = [vp]
    L.dmdgpu_work_counts.argtypes = [vp, _lp, _lp]
    L.dmdgpu_add_instance.argtypes = [vp, C.c_int, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _ip,
                                      C.POINTER(C.c_int)]
    L.dmdgpu_remove_instance.argtypes = [vp, C.c_int]
    L.dmdgpu_instance_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int), _lp, _lp]
    L.dmdgpu_set_budget.argtypes = [vp, C.c_int64]
    L.dmdgpu_bytes_per_site.restype = C.c_int64
    L.dmdgpu_device_bytes.argtypes = [vp, _lp]
    L.dmdgpu_evaluate.argtypes = [vp]
    L.dmdgpu_instance_residual.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp]
    L.dmdgpu_instance_state.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.dmdgpu_instance_set_unknowns.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.dmdgpu_relax_start.argtypes = [vp, C.c_int, C.POINTER(RelaxOpts)]
    L.dmdgpu_relax_tick.argtypes = [vp, C.POINTER(C.c_int)]
    L.dmdgpu_relax_status.argtypes = [vp, C.c_int, C.POINTER(C.c_int), _dp, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]
    L.dmdgpu_relax.argtypes = [vp, C.POINTER(RelaxOpts), C.c_int]
    L.dmdgpu_relax_energy.argtypes = [vp, C.c_int, _dp]
    L.dmdgpu_relax_evals.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.dmdgpu_instance_pairs.argtypes = [vp, C.c_int, _lp, _lp]
    L.dmdgpu_set_graph.argtypes = [vp, C.c_int]
    L.dmdgpu_diffuse_start.argtypes = [vp, C.c_int, C.POINTER(DiffuseOpts)]
    L.dmdgpu_diffusion_rates.argtypes = [vp, C.c_int, C.POINTER(DiffuseOpts), _dp]
    L.dmdgpu_diffusion_neighbors.argtypes = [vp, C.c_int, _ip, _ip]
    L.dmdgpu_jvp.argtypes = [vp, _dp, _dp, _dp]
    L.dmdgpu_jvp_packed.argtypes = [vp, vp, vp, vp, C.c_int]
    L.dmdgpu_diag_blocks.argtypes = [vp, C.c_int, _dp]
    L.dmdgpu_instance_diag_blocks.argtypes = [vp, C.c_int, C.c_int, _dp]
    L.dmdgpu_diag_blocks_noalloc.argtypes = [vp, C.c_int, vp]
    L.dmdgpu_site_DV0_Dq.argtypes = [vp, C.c_int, _ip, _ip, _dp]
    L.dmdgpu_residual_packed.argtypes = [vp, vp, vp, _dp, C.c_int]
    L.dmdgpu_site_DV0_Dxi.argtypes = [vp, C.c_int, _ip, _ip, _dp]
    L.dmdgpu_set_rule.argtypes = [vp, C.c_int]
    L.dmdgpu_set_tuning_precision.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _dn(a):
    return None if a is None else _d(a)


class Context:
    """One device-resident swarm + tables + neighbour lists on one stream."""

    def __init__(self, device=0, table_dir=TABLE_DIR, rc=R_CUTOFF):
        self.L = load_library()
        h = C.c_void_p()
        self._chk(self.L.dmdgpu_create(device, C.byref(h)))
        self.h = h
        for comb in (0, 1, 2):
            self._chk(self.L.dmdgpu_init_adp_MgHx(self.h, comb, table_dir.encode()))
        self._chk(self.L.dmdgpu_set_cutoff(self.h, rc))
        self.n = 0

    def _chk(self, code):
        if code != 0:
            raise DmdError(code, self.L.dmdgpu_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.dmdgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ------------------------------------------------------------
    def set_state(self, sysd):
        n = int(sysd["n"])
        self._keep = {
            "q": np.ascontiguousarray(sysd["q"], dtype=np.float64),
            "sigma": np.ascontiguousarray(sysd["sigma"], dtype=np.float64),
            "x": np.ascontiguousarray(sysd["x"], dtype=np.float64),
            "spc": np.ascontiguousarray(sysd["species"], dtype=np.int32),
            "beta": np.ascontiguousarray(sysd["beta"], dtype=np.float64),
            "gamma": np.ascontiguousarray(sysd["gamma"], dtype=np.float64),
            "box": np.ascontiguousarray(np.concatenate([sysd["lo"], sysd["hi"]]), dtype=np.float64),
            "per": np.ascontiguousarray(sysd["periodic"], dtype=np.int32),
        }
        k = self._keep
        self._chk(self.L.dmdgpu_set_state(self.h, n, _d(k["q"]), _d(k["sigma"]), _d(k["x"]),
                                          _i(k["spc"]), _d(k["beta"]), _d(k["gamma"]),
                                          _d(k["box"]), _i(k["per"])))
        self.n = n

    def update_state(self, q=None, sigma=None, x=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (q, sigma, x)]
        self._chk(self.L.dmdgpu_update_state(self.h, *[_dn(a) for a in arrs]))

    def set_tuning(self, k1=4, k2=4, unroll2=1):
        self._chk(self.L.dmdgpu_set_tuning(self.h, k1, k2, unroll2))

    def set_rule(self, order):
        """Pair/coincident quadrature: 3 (the reference's rule), 5 (27-point
        tensor Gauss-Hermite on the relative coordinate) or 44 (the reference's
        44-point 6D degree-5 pair rule)."""
        self._chk(self.L.dmdgpu_set_rule(self.h, int(order)))

    def set_tuning_precision(self, precision, k1, k2, unroll2):
        """Tuning of one precision ('fp64' or 'split32') only."""
        code = {"fp64": FP64, "split32": SPLIT32}[precision]
        self._chk(self.L.dmdgpu_set_tuning_precision(self.h, code, k1, k2, unroll2))

    def set_precision(self, precision):
        """'fp64' or 'split32' (FP32 points, FP64 sums and totals)."""
        code = {"fp64": FP64, "split32": SPLIT32}[precision]
        self._chk(self.L.dmdgpu_set_precision(self.h, code))

    # -- neighbours ---------------------------------------------------------
    def neighbors(self):
        self._chk(self.L.dmdgpu_neighbors(self.h))

    def get_neighbors(self, cap=MAXNEIGH):
        nn = np.zeros(self.n, dtype=np.int32)
        lst = np.zeros((self.n, cap), dtype=np.int32)
        self._chk(self.L.dmdgpu_get_neighbors(self.h, _i(nn), _i(lst), cap))
        return nn, lst

    # -- evaluation ---------------------------------------------------------
    def mf_rho(self):
        out = np.zeros(self.n)
        self._chk(self.L.dmdgpu_mf_rho(self.h, _d(out)))
        return out

    def site_energies(self):
        V0 = np.zeros(self.n)
        S0 = np.zeros(self.n)
        self._chk(self.L.dmdgpu_site_energies(self.h, _d(V0), _d(S0)))
        return V0, S0

    def residual(self):
        Rq = np.zeros((self.n, 3))
        Rs = np.zeros(self.n)
        Rx = np.zeros(self.n)
        tot = np.zeros(2)
        self._chk(self.L.dmdgpu_residual(self.h, _d(Rq), _d(Rs), _d(Rx), _d(tot)))
        return {"Rq": Rq, "Rs": Rs, "Rx": Rx, "S0_total": tot[0], "V0_total": tot[1]}

    def site_DV0_Dq(self, u, i):
        """[dV0_i/dq_u (3), dV0_i/dsigma_u] per query (u, i): (m, 4)."""
        u = np.ascontiguousarray(np.atleast_1d(u), dtype=np.int32)
        i = np.ascontiguousarray(np.atleast_1d(i), dtype=np.int32)
        out = np.zeros((len(u), 4))
        self._chk(self.L.dmdgpu_site_DV0_Dq(self.h, len(u), _i(u), _i(i), _d(out)))
        return out

    def site_DV0_Dxi(self, u, i):
        """dV0_i/dx_u per query (u, i) in the reference's form: (m,)."""
        u = np.ascontiguousarray(np.atleast_1d(u), dtype=np.int32)
        i = np.ascontiguousarray(np.atleast_1d(i), dtype=np.int32)
        out = np.zeros(len(u))
        self._chk(self.L.dmdgpu_site_DV0_Dxi(self.h, len(u), _i(u), _i(i), _d(out)))
        return out

    def residual_packed(self, X=None, device=None):
        """SNES FormFunction form: X = [q (3n) | sigma (n)] (None: current
        state) -> R = [R_q (3n) | R_sigma (n) | R_x (n)], totals.  With
        device=(dX, dR) device pointers are passed instead (dX may be 0 for
        the current state) and R stays on the device."""
        tot = np.zeros(2)
        if device is not None:
            dX, dR = device
            self._chk(self.L.dmdgpu_residual_packed(self.h, C.c_void_p(dX or None),
                                                    C.c_void_p(dR or None), _d(tot), 1))
            return None, tot
        R = np.zeros(5 * self.n)
        Xc = None if X is None else np.ascontiguousarray(X, dtype=np.float64)
        self._chk(self.L.dmdgpu_residual_packed(self.h, None if Xc is None else Xc.ctypes.data,
                                                R.ctypes.data, _d(tot), 0))
        return R, tot

    def diag_blocks(self, exact=False):
        """Diagonal 3x3 blocks of dR_q/dq, (n, 3, 3), FP64 (see dmdgpu_diag_blocks)."""
        D = np.zeros((self.n, 9))
        self._chk(self.L.dmdgpu_diag_blocks(self.h, 1 if exact else 0, _d(D)))
        return D.reshape(self.n, 3, 3)

    def diag_blocks_noalloc(self, exact, device_ptr):
        """The diagonal blocks into device memory at device_ptr (n x 9 doubles,
        allocated by the caller) without any device allocation by the
        library (dmdgpu_diag_blocks_noalloc)."""
        self._chk(self.L.dmdgpu_diag_blocks_noalloc(self.h, 1 if exact else 0,
                                                    C.c_void_p(device_ptr)))

    def instance_diag_blocks(self, iid, exact=False):
        """Diagonal blocks of a resident instance at its current state, (n, 3, 3)."""
        n = self.instance_info(iid)["n"]
        D = np.zeros((n, 9))
        self._chk(self.L.dmdgpu_instance_diag_blocks(self.h, iid, 1 if exact else 0, _d(D)))
        return D.reshape(n, 3, 3)

    def residual_resident(self):
        self._chk(self.L.dmdgpu_residual_resident(self.h))

    def sync(self):
        self._chk(self.L.dmdgpu_sync(self.h))

    # -- instrumentation ----------------------------------------------------
    def set_timing(self, on):
        self._chk(self.L.dmdgpu_set_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        k = self.L.dmdgpu_kernel_count()
        la = np.zeros(k, dtype=np.int64)
        ms = np.zeros(k)
        self._chk(self.L.dmdgpu_kernel_times(self.h, la.ctypes.data_as(_lp), _d(ms)))
        return {self.L.dmdgpu_kernel_name(i).decode(): (int(la[i]), float(ms[i])) for i in range(k)}

    def reset_kernel_times(self):
        self._chk(self.L.dmdgpu_reset_kernel_times(self.h))

    def work_counts(self):
        p = C.c_int64()
        q = C.c_int64()
        self._chk(self.L.dmdgpu_work_counts(self.h, C.byref(p), C.byref(q)))
        return p.value, q.value

    # -- resident instances ----------------------------------------------------
    def add_instance(self, sysd):
        """Copies a system in, builds its neighbour lists; returns its id."""
        n = int(sysd["n"])
        k = {
            "q": np.ascontiguousarray(sysd["q"], dtype=np.float64),
            "sigma": np.ascontiguousarray(sysd["sigma"], dtype=np.float64),
            "x": np.ascontiguousarray(sysd["x"], dtype=np.float64),
            "spc": np.ascontiguousarray(sysd["species"], dtype=np.int32),
            "beta": np.ascontiguousarray(sysd["beta"], dtype=np.float64),
            "gamma": np.ascontiguousarray(sysd["gamma"], dtype=np.float64),
            "box": np.ascontiguousarray(np.concatenate([sysd["lo"], sysd["hi"]]), dtype=np.float64),
            "per": np.ascontiguousarray(sysd["periodic"], dtype=np.int32),
        }
        iid = C.c_int(-1)
        self._chk(self.L.dmdgpu_add_instance(self.h, n, _d(k["q"]), _d(k["sigma"]), _d(k["x"]),
                                             _i(k["spc"]), _d(k["beta"]), _d(k["gamma"]),
                                             _d(k["box"]), _i(k["per"]), C.byref(iid)))
        return iid.value

    def remove_instance(self, iid):
        self._chk(self.L.dmdgpu_remove_instance(self.h, iid))

    def instance_info(self, iid):
        n = C.c_int()
        p = C.c_int64()
        e = C.c_int64()
        self._chk(self.L.dmdgpu_instance_info(self.h, iid, C.byref(n), C.byref(p), C.byref(e)))
        return {"n": n.value, "pairs": p.value, "arena_sites": e.value}

    def set_budget(self, nbytes):
        self._chk(self.L.dmdgpu_set_budget(self.h, int(nbytes)))

    def bytes_per_site(self):
        return int(self.L.dmdgpu_bytes_per_site())

    def device_bytes(self):
        """Bytes of every device allocation the context holds now."""
        b = C.c_int64()
        self._chk(self.L.dmdgpu_device_bytes(self.h, C.byref(b)))
        return b.value

    def evaluate(self):
        self._chk(self.L.dmdgpu_evaluate(self.h))

    def instance_residual(self, iid):
        n = self.instance_info(iid)["n"]
        Rq = np.zeros((n, 3))
        Rs = np.zeros(n)
        Rx = np.zeros(n)
        tot = np.zeros(2)
        self._chk(self.L.dmdgpu_instance_residual(self.h, iid, _d(Rq), _d(Rs), _d(Rx), _d(tot)))
        return {"Rq": Rq, "Rs": Rs, "Rx": Rx, "S0_total": tot[0], "V0_total": tot[1]}

    def instance_state(self, iid):
        n = self.instance_info(iid)["n"]
        q = np.zeros((n, 3))
        sg = np.zeros(n)
        x = np.zeros(n)
        self._chk(self.L.dmdgpu_instance_state(self.h, iid, _d(q), _d(sg), _d(x)))
        return q, sg, x

    def instance_set_unknowns(self, iid, q=None, sigma=None, x=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (q, sigma, x)]
        self._chk(self.L.dmdgpu_instance_set_unknowns(self.h, iid, *[_dn(a) for a in arrs]))

    # -- relaxation ------------------------------------------------------------
    def relax_start(self, iid, opts=None):
        o = opts if opts is not None else relax_opts()
        self._chk(self.L.dmdgpu_relax_start(self.h, iid, C.byref(o)))

    def relax_tick(self):
        r = C.c_int()
        self._chk(self.L.dmdgpu_relax_tick(self.h, C.byref(r)))
        return r.value

    def relax_status(self, iid):
        st = C.c_int()
        rn = C.c_double()
        nw = C.c_int()
        ev = C.c_int()
        self._chk(self.L.dmdgpu_relax_status(self.h, iid, C.byref(st), C.byref(rn), C.byref(nw),
                                             C.byref(ev)))
        return {"state": RELAX_STATE[st.value], "rnorm": rn.value, "newton": nw.value,
                "evals": ev.value}

    def relax_evals(self, iid):
        """(FP64, split) residual evaluations of the instance's solve."""
        a, b = C.c_int(), C.c_int()
        self._chk(self.L.dmdgpu_relax_evals(self.h, iid, C.byref(a), C.byref(b)))
        return a.value, b.value

    def instance_pairs(self, iid):
        """(active ordered pairs, those with an Mg end) of an instance."""
        a, b = C.c_int64(), C.c_int64()
        self._chk(self.L.dmdgpu_instance_pairs(self.h, iid, C.byref(a), C.byref(b)))
        return a.value, b.value

    def relax_energy(self, iid):
        """Free energy sum_i S0_i at the solve's current iterate."""
        e = C.c_double()
        self._chk(self.L.dmdgpu_relax_energy(self.h, iid, C.byref(e)))
        return e.value

    def relax(self, opts=None, max_ticks=100000):
        o = opts if opts is not None else relax_opts()
        self._chk(self.L.dmdgpu_relax(self.h, C.byref(o), max_ticks))

    def set_graph(self, on):
        self._chk(self.L.dmdgpu_set_graph(self.h, 1 if on else 0))

    def jvp(self, q, sigma, vq, vsigma):
        """J v of the (q, sigma) residual at (q, sigma): returns (Jv_q, Jv_sigma)."""
        n = self.n
        X = np.ascontiguousarray(np.concatenate([np.ravel(q), np.ravel(sigma)]), dtype=np.float64)
        V = np.ascontiguousarray(np.concatenate([np.ravel(vq), np.ravel(vsigma)]), dtype=np.float64)
        JV = np.zeros(4 * n)
        self._chk(self.L.dmdgpu_jvp(self.h, _d(X), _d(V), _d(JV)))
        return JV[:3 * n].reshape(n, 3), JV[3 * n:]

    def jvp_packed(self, X=None, V=None, device=None):
        """dmdgpu_jvp_packed: JV = J V at X, packed [q (3n) | sigma (n)] vectors.
        Host arrays X, V -> JV; or device=(dX, dV, dJV) device pointers."""
        if device is not None:
            dX, dV, dJ = device
            self._chk(self.L.dmdgpu_jvp_packed(self.h, C.c_void_p(dX), C.c_void_p(dV),
                                               C.c_void_p(dJ), 1))
            return None
        Xc = np.ascontiguousarray(X, dtype=np.float64)
        Vc = np.ascontiguousarray(V, dtype=np.float64)
        JV = np.zeros(4 * self.n)
        self._chk(self.L.dmdgpu_jvp_packed(self.h, Xc.ctypes.data, Vc.ctypes.data, JV.ctypes.data,
                                           0))
        return JV

    # -- diffusion step -------------------------------------------------------
    def diffuse_start(self, iid, opts=None):
        o = opts if opts is not None else diffuse_opts()
        self._chk(self.L.dmdgpu_diffuse_start(self.h, iid, C.byref(o)))

    def diffusion_rates(self, iid, opts=None):
        o = opts if opts is not None else diffuse_opts()
        n = self.instance_info(iid)["n"]
        xd = np.zeros(n)
        self._chk(self.L.dmdgpu_diffusion_rates(self.h, iid, C.byref(o), _d(xd)))
        return xd

    def diffusion_neighbors(self, iid):
        n = self.instance_info(iid)["n"]
        cnt = np.zeros(n, dtype=np.int32)
        lst = np.zeros((n, DCAP), dtype=np.int32)
        self._chk(self.L.dmdgpu_diffusion_neighbors(self.h, iid, _i(cnt), _i(lst)))
        return cnt, lst
