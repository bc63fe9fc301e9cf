"""Launch-shape autotuner for batches of registrations of different volume sizes.

A batch (fastreg_amd.Batch) runs every volume's pass A and pass B in one launch each per
iteration.  How many z-slices one thread block marches (the z-chunk) trades halo re-reads and
per-block set-up against load balance across the GPU's 256 CUs, and the best choice for a volume
inside a batch differs from its best choice alone (alone, a volume wants its whole grid resident
at once; in a batch, blocks of all volumes share the CUs).  The tuner chooses, for each distinct
volume size, pass A's z-chunk, and for the batch as a whole the z-chunk of pass B, by timing a few
iterations of the whole batch for each candidate (coordinate descent, largest sizes first).
Results do not depend on the chunking; only the time does.
"""
from __future__ import annotations

CAND_A = (8, 16, 30, 60)
CAND_B = (2, 4, 8)


def tune_batch(batch, iters=3, cand_a=CAND_A, cand_b=CAND_B, log=None):
    """Choose z-chunks for `batch` (a fastreg_amd.Batch whose Solvers are initialised) and apply
    them.  Returns ({dims: (zc_a, zc_b)}, best device ms for `iters` iterations).  The Solvers are
    re-initialised from their inputs afterwards."""
    solvers = batch.solvers
    classes = {}
    for s in solvers:
        classes.setdefault(s.dims, []).append(s)
    order = sorted(classes, key=lambda d: -(d[0] * d[1] * d[2]))
    cur = {d: classes[d][0].chunking() for d in classes}

    def apply():
        for d, ss in classes.items():
            for s in ss:
                s.set_chunking(*cur[d])
        batch.relayout()

    def measure():
        for s in solvers:
            s.reinit()
        batch.time_iterations(1)  # warm-up (caches, first-launch costs)
        for s in solvers:
            s.reinit()
        return batch.time_iterations(iters)

    apply()
    best = measure()
    if log:
        log(f"autotune: start {best:.2f} ms / {iters} it")
    for d in order:
        for za in cand_a:
            old = cur[d]
            if za == old[0]:
                continue
            cur[d] = (za, old[1])
            apply()
            t = measure()
            if t < best:
                best = t
                if log:
                    log(f"autotune: {d} pass A z-chunk {za}: {t:.2f} ms")
            else:
                cur[d] = old
    for z in cand_b:
        old = dict(cur)
        if all(v[1] == z for v in cur.values()):
            continue
        cur = {d: (v[0], z) for d, v in cur.items()}
        apply()
        t = measure()
        if t < best:
            best = t
            if log:
                log(f"autotune: pass B z-chunk {z}: {t:.2f} ms")
        else:
            cur = old
    apply()
    for s in solvers:
        s.reinit()
    return dict(cur), best
