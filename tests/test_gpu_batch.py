"""Batched registrations (fr_batch): several volumes of different sizes advanced together, one
launch per pass per iteration, the whole run replayed from a hipGraph.  Every volume must equal
its own oracle run bit for bit, on every replay of the graph, with eager work (re-initialisation,
single-volume runs) between the replays."""
import sys

import numpy as np
import pytest

import fastreg_amd as fr
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu

DIMS = [(24, 20, 18), (33, 17, 21), (16, 16, 16), (40, 12, 30), (70, 21, 12), (64, 15, 14)]  # 64: one x tile per row


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _volumes(O, iters, verbosity=0, converge=None):
    """Solvers and oracle references for DIMS; volume `converge` gets a huge MAX_UPDATE_NORM (it
    stops after its first iteration while the others run on)."""
    out = []
    for k, dims in enumerate(DIMS):
        X, Y, Z = dims
        pair = phantoms.blobs(shape=(Z, Y, X), seed=100 + k)
        thr = 1e30 if k == converge else -1e-10
        p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=iters, alpha=0.3,
                              max_update_norm=thr, verbosity=verbosity, tau=0.2 + 0.05 * k)
        g = fr.Solver(p)
        g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        o = O.Oracle(dims, "f32")
        o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, p.rho_0)
        r = o.run(O.SolverParams(max_iter=iters, alpha=0.3, max_update_norm=thr, verbosity=verbosity,
                                 no_labels=pair.no_labels, tau=0.2 + 0.05 * k))
        out.append((g, o, r))
    return out


def _check(vols, what):
    for k, (g, o, r) in enumerate(vols):
        st = g.stats()
        assert st.iterations == r["iterations"] and st.stop_reason == r["stop_reason"], (what, k, st.iterations, r)
        mh, eh = g.history()
        n = r["iterations"]
        assert np.array_equal(mh[:n], r["max_hist"][:n]), (what, k)
        gs, os_ = g.get_state(), o.get_state()
        for f in ("v", "rho", "psi", "psi_inv"):
            assert np.array_equal(gs[f], os_[f]), (what, k, f)


def test_batch_matches_oracle_on_every_replay(oracle_mod):
    O = oracle_mod
    vols = _volumes(O, 9, converge=2)
    b = fr.Batch([g for g, _, _ in vols])
    try:
        for rep in range(3):
            say("batch replay", rep)
            if rep:
                for g, _, _ in vols:
                    g.reinit()  # eager work between the replays
            b.run()
            assert b.time_ms() > 0
            _check(vols, f"replay {rep}")
            # a single-volume run of another context between replays
            if rep == 1:
                g1 = fr.Solver(vols[0][0].params)
                pair = phantoms.blobs(shape=tuple(reversed(DIMS[0])), seed=100)
                g1.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
                g1.run()
                g1.stats()
                g1.close()
    finally:
        b.close()


def test_batch_relayout_and_verbose_energy(oracle_mod):
    """Per-volume z-chunkings (re-read with fr_batch_relayout) and the logged energy path."""
    O = oracle_mod
    vols = _volumes(O, 6, verbosity=2)
    b = fr.Batch([g for g, _, _ in vols])
    try:
        for k, (g, _, _) in enumerate(vols):
            g.set_chunking(3 + 2 * k, 1 + (k + 1) % 4)
        b.relayout()
        b.run()
        _check(vols, "relayout")
        for g, o, r in vols:
            _, eh = g.history()
            n = r["iterations"]
            np.testing.assert_allclose(eh[:n], r["energy_hist"][:n], rtol=1e-12)
    finally:
        b.close()


def test_batch_rejects_mixed_schedules():
    p1 = fr.default_params(vol_dims=(8, 8, 8), max_iter=5)
    p2 = fr.default_params(vol_dims=(9, 8, 8), max_iter=6)
    g1, g2 = fr.Solver(p1), fr.Solver(p2)
    with pytest.raises(fr.FastRegError, match="MAX_ITER"):
        fr.Batch([g1, g2])


def test_autotuned_batch_matches_oracle(oracle_mod):
    """The autotuner's timing runs leave the volumes re-initialised; the tuned batch still equals
    the oracle."""
    from fastreg_amd import autotune
    O = oracle_mod
    vols = _volumes(O, 5)
    b = fr.Batch([g for g, _, _ in vols])
    try:
        tuned, ms = autotune.tune_batch(b, iters=2, cand_a=(3, 8), cand_b=(1, 2))
        say("tuned", tuned, ms)
        assert set(tuned) == set(DIMS) and ms > 0
        b.run()
        _check(vols, "tuned")
    finally:
        b.close()


def test_batch_of_degenerate_volumes_matches_oracle(oracle_mod):
    """A batch of the smallest volumes and one-voxel-thick interiors next to a 3-tile-wide one."""
    O = oracle_mod
    dims_list = [(3, 3, 3), (5, 4, 3), (3, 17, 4), (131, 11, 5), (4, 5, 40)]
    vols = []
    for k, dims in enumerate(dims_list):
        X, Y, Z = dims
        pair = phantoms.blobs(shape=(Z, Y, X), seed=200 + k)
        p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=7, alpha=0.3)
        g = fr.Solver(p)
        g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        o = O.Oracle(dims, "f32")
        o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, p.rho_0)
        r = o.run(O.SolverParams(max_iter=7, alpha=0.3, no_labels=pair.no_labels))
        vols.append((g, o, r))
    say("degenerate batch", dims_list)
    b = fr.Batch([g for g, _, _ in vols])
    try:
        b.run()
        _check(vols, "degenerate batch")
    finally:
        b.close()


def test_batch_refuses_to_run_after_a_storage_change(oracle_mod):
    """A batch holds its volumes' buffers and kernels of the storage they had when it was laid out:
    after fr_set_storage on one of them, running or timing the batch fails with FR_ERR_STATE (not a
    run on stale buffers), relayout fails while the storages differ, and once every volume has the
    new storage and the batch is relaid out it runs and matches the oracle in that storage."""
    O = oracle_mod
    vols = []
    for k, dims in enumerate([(20, 18, 16), (35, 12, 22)]):
        X, Y, Z = dims
        pair = phantoms.blobs(shape=(Z, Y, X), seed=60 + k)
        p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=5, alpha=0.3)
        g = fr.Solver(p)
        g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        vols.append((g, pair))
    b = fr.Batch([g for g, _ in vols])
    try:
        vols[1][0].set_storage(fr.STORAGE_BF16)
        for call in (b.run, lambda: b.time_iterations(1)):
            with pytest.raises(fr.FastRegError) as e:
                call()
            assert e.value.code == fr.FR_ERR_STATE
        with pytest.raises(fr.FastRegError) as e:
            b.relayout()
        assert e.value.code == fr.FR_ERR_INVALID
        vols[0][0].set_storage(fr.STORAGE_BF16)
        b.relayout()
        refs = []
        for g, pair in vols:
            g.reinit()
            o = O.Oracle(g.dims, "f32", storage=fr.STORAGE_BF16)
            o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
            refs.append((o, o.run(O.SolverParams(max_iter=5, alpha=0.3, no_labels=pair.no_labels))))
        b.run()
        for (g, _), (o, r) in zip(vols, refs):
            assert g.stats().iterations == r["iterations"] == 5
            gs, os_ = g.get_state(), o.get_state()
            for f in ("v", "rho", "psi", "psi_inv"):
                assert np.array_equal(gs[f], os_[f]), f
    finally:
        b.close()
