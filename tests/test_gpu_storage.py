"""The 16-bit storage forms of v and rho (bf16, fp16; arithmetic in FP32) on the GPU against the
oracle with the same storage rounding: bit-exact, like the FP32 form against the FP32 oracle."""
import sys

import numpy as np
import pytest

import fastreg_amd as fr
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu

STORAGES = [fr.STORAGE_BF16, fr.STORAGE_F16]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(O, dims, seed, storage, **kw):
    X, Y, Z = dims
    pair = phantoms.blobs(shape=(Z, Y, X), seed=seed)
    base = dict(vol_dims=dims, no_labels=pair.no_labels, max_iter=10, verbosity=0)
    base.update(kw)
    p = fr.default_params(**base)
    g = fr.Solver(p)
    g.set_storage(storage)
    assert g.storage() == storage
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o = O.Oracle(dims, "f32", storage=storage)
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, p.rho_0)
    return g, o, pair


def _same_state(g, o, what):
    gs, os_ = g.get_state(), o.get_state()
    for k in ("v", "rho", "psi", "psi_inv"):
        a, b = gs[k].view(np.uint32), os_[k].view(np.uint32)   # raw bits: zero signs count
        if not np.array_equal(a, b):
            bad = np.flatnonzero((a != b).reshape(len(a), -1).any(-1) if a.ndim > 1 else (a != b))
            raise AssertionError(f"{what} {k}: {len(bad)} voxels differ, first {bad[:4]}")


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("dims,seed,amp", [((24, 20, 18), 3, 1.5), ((67, 9, 11), 4, 1.5), ((70, 40, 36), 31, 4.0),
                                            ((64, 21, 13), 41, 0.6)])
def test_step_teacher_forced(oracle_mod, storage, dims, seed, amp):
    O = oracle_mod
    say("storage step", storage, dims, amp)
    g, o, pair = _both(O, dims, seed, storage)
    rng = np.random.default_rng(seed)
    st = phantoms.random_state(dims, rng, amp=amp)
    g.set_state(**st)
    o.set_state(**st)
    _same_state(g, o, "set_state")  # both round v and rho the same way
    for t in (0.004, 0.3):
        mg, ag = g.step(t)
        mo, ao = o.step(t)
        assert mg == mo and ag == ao, (mg, mo, ag, ao)
        _same_state(g, o, f"step t={t}")


@pytest.mark.parametrize("storage", STORAGES)
def test_run_with_log(oracle_mod, storage):
    O = oracle_mod
    dims = (32, 28, 24)
    say("storage run", storage)
    g, o, pair = _both(O, dims, 7, storage, max_iter=25, verbosity=2, alpha=0.4)
    g.run()
    r = o.run(O.SolverParams(max_iter=25, verbosity=2, no_labels=pair.no_labels, alpha=0.4))
    st = g.stats()
    assert st.iterations == r["iterations"] == 25
    mh, eh = g.history()
    assert np.array_equal(mh, r["max_hist"])
    np.testing.assert_allclose(eh, r["energy_hist"][:25], rtol=1e-12)
    _same_state(g, o, "run")
    for a, b, nm in zip(g.finalize(), o.outputs(), ("img1", "lab1", "img2", "lab2")):
        assert np.array_equal(a, b), nm


@pytest.mark.parametrize("storage", STORAGES)
def test_kept_state_in_16_bit_storage(oracle_mod, storage):
    """A context that keeps its state: the second registration resets rho inside the (v, rho)
    records and keeps v; equal to the oracle with the same storage doing the same."""
    O = oracle_mod
    dims = (28, 22, 20)
    say("kept state", storage)
    g, o, pair = _both(O, dims, 71, storage, max_iter=12, alpha=0.3)
    g.set_keep_state(True)
    prm = O.SolverParams(max_iter=12, alpha=0.3, no_labels=pair.no_labels)
    for k in range(2):
        g.register(pair.target, pair.target_labels, pair.source, pair.source_labels)
        st = o.get_state()
        if k:
            o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
            o.set_state(v=st["v"], psi=st["psi"], psi_inv=st["psi_inv"])
        o.run(prm)
        _same_state(g, o, f"registration {k}")


def test_batch_in_bf16(oracle_mod):
    O = oracle_mod
    vols = []
    for k, dims in enumerate([(20, 18, 16), (35, 12, 22)]):
        g, o, pair = _both(O, dims, 40 + k, fr.STORAGE_BF16, max_iter=6, alpha=0.3)
        r = o.run(O.SolverParams(max_iter=6, alpha=0.3, no_labels=pair.no_labels))
        vols.append((g, o, r))
    b = fr.Batch([g for g, _, _ in vols])
    try:
        b.run()
        for g, o, r in vols:
            assert g.stats().iterations == r["iterations"]
            _same_state(g, o, "batch")
    finally:
        b.close()
