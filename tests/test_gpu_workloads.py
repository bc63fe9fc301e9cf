"""GPU parity at the benchmark workloads' own shapes and weights, against the oracle.

- the 64^3 blob pair (9 labels) with the params.ini weights for its full 200 iterations, energy
  logged at iterations 1, 100 and 200;
- the 182x218x182 brain phantom (33 labels) with the params.ini weights for its full 300
  iterations (the headline benchmark's registration), and teacher-forced steps there;
- the 64^3 pair with v and rho stored in BF16 and in FP16, 300 iterations, against the oracle
  that rounds the same way;
- a batch holding one volume of each of the five sizes 48^3, 96^3, 128x160x128, 160x192x160 and
  182x218x182, launch shapes from the autotuner, in FP32 and BF16 storage.

Bit-exact is the bar, as in the other parity tests: every float field, label, max norm, Dice
count and min det J equals the oracle's.
"""
import sys

import numpy as np
import pytest

import fastreg_amd as fr
from fastreg_amd import autotune, phantoms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same_state(g, o, what):
    gs, os_ = g.get_state(), o.get_state()
    for k in ("v", "rho", "psi", "psi_inv"):
        a, b = gs[k].view(np.uint32), os_[k].view(np.uint32)   # raw bits: zero signs count
        if not np.array_equal(a, b):
            bad = np.flatnonzero((a != b).reshape(len(a), -1).any(-1) if a.ndim > 1 else (a != b))
            raise AssertionError(f"{what} {k}: {len(bad)} voxels differ, first {bad[:4]}")


def _same_outputs(g, o, no_labels, what):
    for a, b, nm in zip(g.finalize(), o.outputs(), ("img1", "lab1", "img2", "lab2")):
        assert np.array_equal(a.reshape(-1), b), (what, nm)
    _, ov, vx = g.dice()
    oov, ovx, bad = o.dice_counts(no_labels)
    assert bad == 0 and np.array_equal(ov, oov) and np.array_equal(vx, ovx), what
    st = g.stats()
    assert (st.min_det_true, st.min_det_ref) == o.min_det(), what
    return st


@pytest.fixture(scope="module")
def brain():
    return phantoms.brain()


def test_c1_blob_pair_full_200_iterations(oracle_mod):
    """fr_register on the 64^3 pair, params.ini weights, MAX_ITER 200, --verbose energies."""
    O = oracle_mod
    pair = phantoms.blobs()
    dims = pair.dims
    say("64^3, 200 iterations", dims)
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=200, verbosity=1)
    assert (p.alpha, p.tau, p.w_reg, p.rho_0, p.p, p.C) == tuple(np.float32(v) for v in (0.004, 0.2, 0.2, 1, 2, 0.25))
    g = fr.Solver(p)
    outs = g.register(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    r = o.run(O.SolverParams(max_iter=200, verbosity=1, no_labels=pair.no_labels))
    for a, b, nm in zip(outs, o.outputs(), ("img1", "lab1", "img2", "lab2")):
        assert np.array_equal(a.reshape(-1), b), nm
    st = g.stats()
    assert st.iterations == r["iterations"] == 200
    mh, eh = g.history()
    assert np.array_equal(mh, r["max_hist"])
    logged = np.isfinite(r["energy_hist"][:200])
    assert np.array_equal(np.isfinite(eh), logged) and logged.sum() == 3
    np.testing.assert_allclose(eh[logged], r["energy_hist"][:200][logged], rtol=1e-12)
    _, ov, vx = g.dice()
    oov, ovx, _ = o.dice_counts(pair.no_labels)
    assert np.array_equal(ov, oov) and np.array_equal(vx, ovx)
    assert (st.min_det_true, st.min_det_ref) == o.min_det()
    say("  dice", st.dice_all, "min det", st.min_det_true)


def test_c2_brain_steps_at_the_reference_weights(oracle_mod, brain):
    """Teacher-forced iterations at 182x218x182 with the params.ini weights (alpha 0.004) from a
    random smooth state displaced by up to ~2 voxels, with t of iterations 1, 2 and 150."""
    O = oracle_mod
    pair = brain
    dims = pair.dims
    say("182x218x182 steps", dims)
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=3)
    g = fr.Solver(p)
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    st = phantoms.random_state(dims, np.random.default_rng(2), amp=2.0)
    g.set_state(**st)
    o.set_state(**st)
    for t in (0.004, 0.008, float(np.float32(0.004) * 150)):
        mg, ag = g.step(t)
        mo, ao = o.step(t)
        assert mg == mo and ag == ao, (t, mg, mo, ag, ao)
    _same_state(g, o, "brain steps")


def test_c2_brain_full_300_iterations(oracle_mod, brain):
    """The headline benchmark's registration: 182x218x182, 33 labels, params.ini weights, 300
    iterations.  History, final state, the four outputs, Dice counts and min det J equal the
    oracle's."""
    O = oracle_mod
    pair = brain
    dims = pair.dims
    say("182x218x182, 300 iterations", dims)
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=300)
    g = fr.Solver(p)
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    g.run()
    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    r = o.run(O.SolverParams(max_iter=300, no_labels=pair.no_labels))
    st = g.stats()
    assert st.iterations == r["iterations"] == 300
    assert np.array_equal(g.history()[0], r["max_hist"])
    _same_state(g, o, "brain 300")
    st = _same_outputs(g, o, pair.no_labels, "brain 300")
    say("  dice", st.dice_all, "min det", st.min_det_true)


@pytest.mark.parametrize("storage", [fr.STORAGE_BF16, fr.STORAGE_F16])
def test_c3_blob_pair_16bit_storage_300_iterations(oracle_mod, storage):
    """The 64^3 pair with v and rho in BF16 / FP16 (FP32 arithmetic): teacher-forced steps from a
    random state, then 300 iterations with the params.ini weights and every energy logged, against
    the oracle with the same rounding."""
    O = oracle_mod
    pair = phantoms.blobs()
    dims = pair.dims
    say("64^3 16-bit storage", storage)
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=300, verbosity=2)
    g = fr.Solver(p)
    g.set_storage(storage)
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o = O.Oracle(dims, "f32", storage=storage)
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    st = phantoms.random_state(dims, np.random.default_rng(3), amp=1.5)
    g.set_state(**st)
    o.set_state(**st)
    for t in (0.004, 0.3):
        mg, ag = g.step(t)
        mo, ao = o.step(t)
        assert mg == mo and ag == ao, (t, mg, mo, ag, ao)
    _same_state(g, o, "16-bit steps")
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    g.run()
    r = o.run(O.SolverParams(max_iter=300, verbosity=2, no_labels=pair.no_labels))
    s = g.stats()
    assert s.iterations == r["iterations"] == 300
    mh, eh = g.history()
    assert np.array_equal(mh, r["max_hist"])
    np.testing.assert_allclose(eh, r["energy_hist"][:300], rtol=1e-12)
    _same_state(g, o, "16-bit run")
    s = _same_outputs(g, o, pair.no_labels, "16-bit run")
    say("  dice", s.dice_all, "min det", s.min_det_true)


@pytest.mark.parametrize("storage", [fr.STORAGE_F32, fr.STORAGE_BF16])
def test_c4_batch_of_the_five_sizes(oracle_mod, storage):
    """One batch holding a 48^3, 96^3, 128x160x128, 160x192x160 and 182x218x182 blob pair (seeds
    1000+i, as the batched workload generates them), z-chunks chosen by the autotuner, 6
    iterations with the params.ini weights: every volume equals its own oracle run."""
    O = oracle_mod
    say("batch of five sizes", storage)
    vols = []
    for k, dims in enumerate([(48, 48, 48), (96, 96, 96), (128, 160, 128), (160, 192, 160), (182, 218, 182)]):
        X, Y, Z = dims
        pair = phantoms.blobs(shape=(Z, Y, X), seed=1000 + k)
        p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=6)
        g = fr.Solver(p)
        g.set_storage(storage)
        g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        o = O.Oracle(dims, "f32", storage=storage)
        o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        r = o.run(O.SolverParams(max_iter=6, no_labels=pair.no_labels))
        vols.append((g, o, r, pair))
    b = fr.Batch([v[0] for v in vols])
    try:
        tuned, _ = autotune.tune_batch(b, iters=1)
        say("  z-chunks", tuned)
        b.run()
        for k, (g, o, r, pair) in enumerate(vols):
            st = g.stats()
            assert st.iterations == r["iterations"] == 6, k
            assert np.array_equal(g.history()[0], r["max_hist"]), k
            _same_state(g, o, f"batch volume {k}")
            _same_outputs(g, o, pair.no_labels, f"batch volume {k}")
    finally:
        b.close()
