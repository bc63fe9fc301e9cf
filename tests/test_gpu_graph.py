"""The iteration loop inside graphs, against the oracle.

- fr_solver_run replays its own captured hipGraph: the same context run again and again from both
  ping-pong parities (an odd MAX_ITER flips the parity each run) gives the oracle's results every
  time.
- A caller captures fr_solver_run into its own graph (fr_set_stream + hipStreamBeginCapture) and
  replays it several times, each replay after an eager fr_solver_init and followed by eager work
  (statistics, state reads, fr_solver_finalize): every replay's history, energies, state, outputs
  and Dice equal the oracle's, and an eager fr_solver_run afterwards still does.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import fastreg_amd as fr
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu

HIP_CAPTURE_THREAD_LOCAL = 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _hip():
    h = C.CDLL("libamdhip64.so")
    for name in ("hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize", "hipStreamBeginCapture",
                 "hipStreamEndCapture", "hipGraphInstantiate", "hipGraphLaunch", "hipGraphExecDestroy",
                 "hipGraphDestroy"):
        getattr(h, name).restype = C.c_int
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
    h.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    h.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t]
    h.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]
    h.hipGraphExecDestroy.argtypes = [C.c_void_p]
    h.hipGraphDestroy.argtypes = [C.c_void_p]
    return h


def _ok(rc, what):
    assert rc == 0, f"{what}: hipError {rc}"


def _reference(O, pair, iters):
    o = O.Oracle(pair.dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    r = o.run(O.SolverParams(max_iter=iters, verbosity=2, no_labels=pair.no_labels))
    st = o.get_state()
    outs = o.outputs()
    ov, vx, bad = o.dice_counts(pair.no_labels)
    md = o.min_det()
    o.close()
    return r, st, outs, (ov, vx, bad), md


def _check(g, ref, iters, pair, what):
    r, st, outs, (ov, vx, bad), md = ref
    s = g.stats()
    assert s.iterations == r["iterations"] == iters, what
    mh, eh = g.history()
    assert np.array_equal(mh, r["max_hist"]), what
    np.testing.assert_allclose(eh, r["energy_hist"][:iters], rtol=1e-12, err_msg=what)
    gs = g.get_state()
    for k in ("v", "rho", "psi", "psi_inv"):
        assert np.array_equal(gs[k].view(np.uint32), st[k].view(np.uint32)), (what, k)
    for a, b, nm in zip(g.finalize(), outs, ("img1", "lab1", "img2", "lab2")):
        assert np.array_equal(a.reshape(-1), b), (what, nm)
    _, gov, gvx = g.dice()
    assert bad == 0 and np.array_equal(gov, ov) and np.array_equal(gvx, vx), what
    s = g.stats()
    assert (s.min_det_true, s.min_det_ref) == md, what


def test_own_graph_replays_from_both_parities(oracle_mod):
    """Four runs in a row on one context, each continuing from the state the last one left: an odd
    MAX_ITER starts them at parities 0, 1, 0, 1, so both captured loops are made and replayed.
    Each run's history and state equal the oracle's, run the same way."""
    O = oracle_mod
    pair = phantoms.blobs(shape=(30, 34, 40), seed=21)
    iters = 11
    o = O.Oracle(pair.dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    g = fr.Solver(fr.default_params(vol_dims=pair.dims, no_labels=pair.no_labels, max_iter=iters, verbosity=2))
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    for rep in range(4):
        say("own graph, run", rep)
        g.run()
        r = o.run(O.SolverParams(max_iter=iters, verbosity=2, no_labels=pair.no_labels))
        assert g.stats().iterations == r["iterations"] == iters, rep
        mh, eh = g.history()
        assert np.array_equal(mh, r["max_hist"]), rep
        np.testing.assert_allclose(eh, r["energy_hist"][:iters], rtol=1e-12, err_msg=f"run {rep}")
        gs, st = g.get_state(), o.get_state()
        for k in ("v", "rho", "psi", "psi_inv"):
            assert np.array_equal(gs[k].view(np.uint32), st[k].view(np.uint32)), (rep, k)
    o.close()


def test_caller_captured_loop_matches_oracle_on_every_replay(oracle_mod):
    hip = _hip()
    pair = phantoms.blobs(shape=(30, 34, 40), seed=22)
    iters = 9
    ref = _reference(oracle_mod, pair, iters)
    g = fr.Solver(fr.default_params(vol_dims=pair.dims, no_labels=pair.no_labels, max_iter=iters, verbosity=2))
    stream = C.c_void_p()
    _ok(hip.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    graph, exe = C.c_void_p(), C.c_void_p()
    try:
        fr.check(fr.lib().fr_set_stream(g.h, stream))
        g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        _ok(hip.hipStreamSynchronize(stream), "sync")
        say("caller capture of fr_solver_run")
        _ok(hip.hipStreamBeginCapture(stream, HIP_CAPTURE_THREAD_LOCAL), "hipStreamBeginCapture")
        rc = fr.lib().fr_solver_run(g.h)
        rc_end = hip.hipStreamEndCapture(stream, C.byref(graph))
        fr.check(rc)
        _ok(rc_end, "hipStreamEndCapture")
        _ok(hip.hipGraphInstantiate(C.byref(exe), graph, None, None, 0), "hipGraphInstantiate")
        for rep in range(3):
            say("replay", rep)
            g.reinit()  # eager, outside the capture
            _ok(hip.hipGraphLaunch(exe, stream), "hipGraphLaunch")
            _ok(hip.hipStreamSynchronize(stream), "sync")
            _check(g, ref, iters, pair, f"replay {rep}")
        say("eager run after the replays")
        g.reinit()
        g.run()
        _check(g, ref, iters, pair, "eager run")
    finally:
        if exe.value:
            hip.hipGraphExecDestroy(exe)
        if graph.value:
            hip.hipGraphDestroy(graph)
        fr.lib().fr_set_stream(g.h, None)
        g.close()
        hip.hipStreamDestroy(stream)


def test_kept_state_init_inside_a_capture_needs_the_run_collected():
    """A context that keeps its state must read back an earlier run before its next fr_solver_init;
    inside a stream capture that read cannot run, so the call returns FR_ERR_STATE (and captures
    nothing).  Outside the capture, after the run is collected, it succeeds."""
    hip = _hip()
    pair = phantoms.blobs(shape=(16, 18, 20), seed=23)
    g = fr.Solver(fr.default_params(vol_dims=pair.dims, no_labels=pair.no_labels, max_iter=4))
    stream = C.c_void_p()
    _ok(hip.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    graph = C.c_void_p()
    try:
        fr.check(fr.lib().fr_set_stream(g.h, stream))
        g.set_keep_state(True)
        g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
        fr.check(fr.lib().fr_solver_run(g.h))  # not collected
        _ok(hip.hipStreamSynchronize(stream), "sync")
        inputs = [b.ptr for b in g._inputs]
        _ok(hip.hipStreamBeginCapture(stream, HIP_CAPTURE_THREAD_LOCAL), "hipStreamBeginCapture")
        rc = fr.lib().fr_solver_init(g.h, *inputs)
        rc_end = hip.hipStreamEndCapture(stream, C.byref(graph))
        assert rc == fr.FR_ERR_STATE, rc
        _ok(rc_end, "hipStreamEndCapture")
        assert g.stats().iterations == 4  # collects the run
        fr.check(fr.lib().fr_solver_init(g.h, *inputs))
        g.run()
        assert g.stats().iterations == 4
    finally:
        if graph.value:
            hip.hipGraphDestroy(graph)
        fr.lib().fr_set_stream(g.h, None)
        g.close()
        hip.hipStreamDestroy(stream)
