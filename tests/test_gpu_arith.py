"""The contracted arithmetic (fr_set_arith FR_ARITH_FMA): the reference's multiply-adds fused and
the density's division by a reciprocal.  Its results are not the reference order's bits; the bar
is the path's tolerance, a relative L2 deviation of at most 1e-5 per iteration in every float
field against the oracle (the reference's FP32 order), while warps at a fixed deformation stay
bit-exact (they do not depend on the mode).  The API refuses the combinations the mode does not
support (16-bit storage, batches)."""
import sys

import numpy as np
import pytest

import fastreg_amd as fr
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L2 per iteration, every float field


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(O, dims, seed, **kw):
    X, Y, Z = dims
    pair = phantoms.blobs(shape=(Z, Y, X), seed=seed)
    base = dict(vol_dims=dims, no_labels=pair.no_labels, max_iter=10, verbosity=0)
    base.update(kw)
    p = fr.default_params(**base)
    g = fr.Solver(p)
    g.set_arith(fr.ARITH_FMA)
    assert g.arith() == fr.ARITH_FMA
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, p.rho_0)
    return g, o, pair


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _within(g, o, what):
    gs, os_ = g.get_state(), o.get_state()
    errs = {k: _rel(gs[k], os_[k]) for k in ("v", "rho", "psi", "psi_inv")}
    assert all(e <= TOL for e in errs.values()), (what, errs)
    return errs


@pytest.mark.parametrize("dims,seed,amp", [((24, 20, 18), 3, 1.5), ((131, 19, 25), 32, 1.5), ((64, 21, 13), 41, 0.6),
                                           ((70, 40, 36), 31, 4.0)])
def test_fma_steps_within_tolerance(oracle_mod, dims, seed, amp):
    """Teacher-forced steps from the same random state: every field within 1e-5 relative L2 of the
    oracle after each step (three x tiles; one tile per row at 64 wide; displacements that send
    voxels to the recovery after the march), and the steps do differ from the reference order."""
    O = oracle_mod
    say("fma step", dims, amp)
    g, o, pair = _both(O, dims, seed)
    rng = np.random.default_rng(seed)
    st = phantoms.random_state(dims, rng, amp=amp)
    g.set_state(**st)
    o.set_state(**st)
    differs = False
    for t in (0.004, 0.3):
        mg, _ = g.step(t)
        mo, _ = o.step(t)
        assert abs(mg - mo) <= 1e-5 * mo, (mg, mo)
        errs = _within(g, o, f"t={t}")
        differs |= errs["v"] > 0
        g.set_state(**o.get_state())  # teacher forcing: the next step starts from the oracle's state
    assert differs, "the contracted arithmetic gave the reference order's bits"


def test_fma_run_tracks_the_oracle(oracle_mod):
    """A 25-iteration run (alpha 0.4, energy logged): the max-norm history within 1e-4 relative,
    the final fields within 1e-4, Dice within 1e-3 of the oracle's."""
    O = oracle_mod
    dims = (32, 28, 24)
    g, o, pair = _both(O, dims, 7, max_iter=25, verbosity=2, alpha=0.4)
    g.run()
    r = o.run(O.SolverParams(max_iter=25, verbosity=2, no_labels=pair.no_labels, alpha=0.4))
    assert g.stats().iterations == r["iterations"] == 25
    mh, eh = g.history()
    np.testing.assert_allclose(mh, r["max_hist"], rtol=1e-4)
    np.testing.assert_allclose(eh, r["energy_hist"][:25], rtol=1e-4)
    gs, os_ = g.get_state(), o.get_state()
    for k in ("v", "rho", "psi", "psi_inv"):
        assert _rel(gs[k], os_[k]) <= 1e-4, k
    g.finalize()
    o.outputs()
    dice, ov, vx = g.dice()
    oov, ovx, _ = o.dice_counts(pair.no_labels)
    od = 2 * oov[1:].sum() / ovx[1:].sum()
    gd = 2 * ov[1:].sum() / vx[1:].sum()
    assert abs(gd - od) <= 1e-3, (gd, od)


def test_fma_warps_at_a_fixed_state_are_bit_exact(oracle_mod):
    """The final warps of a given state do not depend on the arithmetic mode: bit-exact."""
    O = oracle_mod
    dims = (40, 22, 18)
    g, o, pair = _both(O, dims, 33)
    st = phantoms.random_state(dims, np.random.default_rng(33), amp=2.0)
    g.set_state(**st)
    img1, lab1, img2, lab2 = g.finalize()
    for img, lab, vol, labs, psi in ((img1, lab1, pair.target, pair.target_labels, st["psi_inv"]),
                                     (img2, lab2, pair.source, pair.source_labels, st["psi"])):
        phi = np.stack([vol.reshape(-1), labs.reshape(-1).astype(np.float32)], 1)
        e = O.apply(dims, phi, psi)
        assert np.array_equal(np.asarray(img).reshape(-1), e[:, 0])
        assert np.array_equal(np.asarray(lab).reshape(-1), e[:, 1].astype(np.int16))


def test_fma_api_rules():
    """FMA needs FP32 storage both ways round; batches refuse FMA contexts, and a batch laid out
    before a context switched to FMA refuses to run until it is laid out again (which fails while
    the context stays in FMA); switching back is accepted."""
    dims = (20, 18, 16)
    pair = phantoms.blobs(shape=(16, 18, 20), seed=5)
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=3)
    a, b = fr.Solver(p), fr.Solver(p)
    L = fr.lib()
    try:
        assert L.fr_set_arith(a.h, 7) == fr.FR_ERR_INVALID
        a.set_storage(fr.STORAGE_BF16)
        assert L.fr_set_arith(a.h, fr.ARITH_FMA) == fr.FR_ERR_INVALID
        a.set_storage(fr.STORAGE_F32)
        a.set_arith(fr.ARITH_FMA)
        assert L.fr_set_storage(a.h, fr.STORAGE_F16) == fr.FR_ERR_INVALID
        with pytest.raises(fr.FastRegError):
            fr.Batch([a, b])
        a.set_arith(fr.ARITH_REFERENCE)
        for s in (a, b):
            s.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        bt = fr.Batch([a, b])
        try:
            a.set_arith(fr.ARITH_FMA)
            assert L.fr_batch_run(bt.h) == fr.FR_ERR_STATE
            assert L.fr_batch_relayout(bt.h) == fr.FR_ERR_INVALID
            a.set_arith(fr.ARITH_REFERENCE)
            assert L.fr_batch_relayout(bt.h) == fr.FR_OK
            bt.run()
        finally:
            bt.close()
    finally:
        a.close()
        b.close()
