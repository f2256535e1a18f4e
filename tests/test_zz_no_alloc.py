"""GPU test of the allocation-free entry point dmdgpu_diag_blocks_noalloc.

Through the C ABI, against dmdgpu_diag_blocks (the entry point it is a
variant of): bit for bit the same blocks, in the caller's site order, on two
calls in a row -- the second with a smaller system -- in both assembly modes.
The call allocates nothing: the context's device allocations are the same
before and after it, and so is the device's free memory (after the first call
of the process, which may load kernel code).  An instance larger than the
scratch array is refused, and so are tables loaded since the last evaluation
(their upload may allocate) until an allocating call has used them.  Device
memory for the output comes from the same
HIP runtime the library uses (hipMalloc through ctypes), not from the library.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import dmdgpu
import dmdgpu.lattice as lat

pytestmark = pytest.mark.gpu
ERR_ARG, ERR_STATE = 1, 8
SCRATCH_SITES = 65536


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    def __init__(self):
        dmdgpu.load_library()
        self.h = C.CDLL("libamdhip64.so.7")

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p.value

    def free(self, p):
        assert self.h.hipFree(C.c_void_p(p)) == 0

    def to_host(self, p, n):
        out = np.empty(n, dtype=np.float64)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p),
                                C.c_size_t(8 * n), 2) == 0  # hipMemcpyDeviceToHost
        return out

    def free_bytes(self):
        f, t = C.c_size_t(), C.c_size_t()
        assert self.h.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value


def test_diag_blocks_noalloc_matches_allocating_entry_point():
    hip = Hip()
    ctx = dmdgpu.Context(0)
    try:
        systems = [lat.hcp_mgh(20, 15, 15, x_h=(0.0, 0.1), seed=2024),   # 54,000 sites
                   lat.cluster(3, 2, 2, x_h=(0.05, 0.3), seed=7)]       # 144 sites
        first = True
        for sd in systems:
            n = int(sd["n"])
            ctx.set_state(sd)
            ctx.neighbors()
            buf = hip.malloc(8 * 9 * n)
            try:
                for exact in (False, True):
                    say(f"[gpu] noalloc diag blocks, {n} sites, exact={exact}")
                    ref = ctx.diag_blocks(exact).reshape(-1)
                    held = ctx.device_bytes()
                    free0 = hip.free_bytes()
                    ctx.diag_blocks_noalloc(exact, buf)
                    assert ctx.device_bytes() == held
                    if not first:
                        assert hip.free_bytes() == free0
                    first = False
                    got = hip.to_host(buf, 9 * n)
                    assert np.array_equal(got, ref)
            finally:
                hip.free(buf)
        # tables loaded again: refused until an allocating call uploads them
        n = int(systems[-1]["n"])
        buf = hip.malloc(8 * 9 * n)
        try:
            say("[gpu] noalloc diag blocks after a table reload")
            for comb in (0, 1, 2):
                assert ctx.L.dmdgpu_init_adp_MgHx(ctx.h, comb, dmdgpu.TABLE_DIR.encode()) == 0
            with pytest.raises(dmdgpu.DmdError) as e:
                ctx.diag_blocks_noalloc(False, buf)
            assert e.value.code == ERR_STATE
            ref = ctx.diag_blocks(False).reshape(-1)
            ctx.diag_blocks_noalloc(False, buf)
            assert np.array_equal(hip.to_host(buf, 9 * n), ref)
        finally:
            hip.free(buf)
        big = lat.hcp_mgh(24, 18, 18, x_h=(0.0, 0.1), seed=5)
        assert big["n"] > SCRATCH_SITES
        ctx.set_state(big)
        ctx.neighbors()
        buf = hip.malloc(8)
        try:
            with pytest.raises(dmdgpu.DmdError) as e:
                ctx.diag_blocks_noalloc(False, buf)
            assert e.value.code == ERR_ARG
        finally:
            hip.free(buf)
    finally:
        ctx.close()
