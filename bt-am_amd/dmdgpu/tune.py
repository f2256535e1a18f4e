"""Autotuner for the phase kernels.

Tunables, per precision: lanes of a wavefront per site in phase 1 (k1),
blocks per 256-site chunk of the pair kernels (k2), and whether phase 2
advances two or three points per loop iteration (unroll2).  Each
candidate is timed with the context's per-kernel HIP events on the current
state and lists; phase 1 and phase 2 are independent kernels, so each picks its
own best setting.
"""
LANES = (4, 8, 16)
BLOCKS = (4, 8, 16, 32, 64)  # phase-2 blocks per chunk


def autotune(ctx, precision="fp64", reps=3, lanes=LANES, blocks=BLOCKS, unrolls=(0, 1)):
    """Tunes one precision ('fp64' or 'split32') on the context's resident
    state; returns ((k1, k2, unroll2), table) and leaves that precision tuned
    and selected.  table maps (k1, k2, unroll2) -> (phase1 ms, phase2 ms)."""
    table = {}
    ctx.set_precision(precision)
    ctx.set_timing(True)
    # phase 1 depends on k1 only and phase 2 on (k2, unroll2): one sweep each
    cands = [(k1, blocks[-2], 1) for k1 in lanes] + [(lanes[-2], k2, un) for k2 in blocks
                                                     for un in unrolls]
    for k1, k2, un in dict.fromkeys(cands):
        ctx.set_tuning_precision(precision, k1, k2, un)
        ctx.residual_resident()  # first launch of this instantiation, untimed
        ctx.sync()
        ctx.reset_kernel_times()
        for _ in range(reps):
            ctx.residual_resident()
        kt = ctx.kernel_times()
        p1 = kt["phase1_site_sums"][1] / max(kt["phase1_site_sums"][0], 1)
        p2 = kt["phase2_gradient"][1] / max(kt["phase2_gradient"][0], 1)
        table[(k1, k2, un)] = (p1, p2)
    ctx.set_timing(False)
    best1 = min(((v[0], k[0]) for k, v in table.items()))[1]
    best2 = min(((v[1], (k[1], k[2])) for k, v in table.items()))[1]
    cfg = (best1, best2[0], best2[1])
    ctx.set_tuning_precision(precision, *cfg)
    ctx.reset_kernel_times()
    return cfg, table
