"""GPU parity tests: the HIP solver (through the C-ABI) against the oracle.

Bit-exact is the bar throughout: the kernels and the oracle perform the same IEEE operations
in the same order (no contraction, fma only where the reference has one, correctly rounded
division and sqrt), so every float field, label and max-norm must be identical.  Energies are
FP64 sums in different orders and are compared with rtol 1e-12.
"""
import sys

import numpy as np
import pytest

import fastreg_amd as fr
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair_dims(pair):
    return pair.dims


def _params(dims, **kw):
    base = dict(vol_dims=dims, no_labels=9, max_iter=10, verbosity=0)
    base.update(kw)
    return fr.default_params(**base)


def test_device_is_gfx950():
    info = fr.device_info()
    say("device:", info)
    assert "gfx950" in info


@pytest.mark.parametrize("dims,seed", [((23, 19, 17), 0), ((64, 5, 9), 1), ((7, 33, 12), 2)])
def test_warp_bit_exact(oracle_mod, dims, seed):
    say("warp", dims, seed)
    O = oracle_mod
    X, Y, Z = dims
    n = X * Y * Z
    rng = np.random.default_rng(seed)
    img = rng.random(n).astype(np.float32)
    lab = rng.integers(0, 50, n).astype(np.int16)
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    psi = np.zeros((n, 4), np.float32)
    psi[:, 0] = xx.ravel() + rng.normal(0, 3, n)
    psi[:, 1] = yy.ravel() + rng.normal(0, 3, n)
    psi[:, 2] = zz.ravel() + rng.normal(0, 3, n)
    psi[::7, :3] = np.round(psi[::7, :3])          # exact integers hit the weight-0 neighbour
    psi[::11, 0] = X - 1                           # clamp edge
    psi[::13, 1] = -5.0                            # below the grid
    psi[5::17, 0] = np.nan                         # non-finite coordinates clamp as fminf(fmaxf(0, p), X-1)
    psi[6::19, 1] = np.inf
    psi[7::23, 2] = -np.inf
    psi[8::29, 0] = 3.0e30
    g_img, g_lab = fr.warp(img, lab, psi, dims)
    phi2 = np.stack([img, lab.astype(np.float32)], 1)
    o = O.apply(dims, phi2, psi)
    assert np.array_equal(g_img, o[:, 0])
    assert np.array_equal(g_lab, o[:, 1].astype(np.int16))


def _both(dims, O, seed, src_pair=None, **kw):
    rng = np.random.default_rng(seed)
    X, Y, Z = dims
    if src_pair is None:
        pair = phantoms.blobs(shape=(Z, Y, X), seed=seed)
    else:
        pair = src_pair
    p = _params(dims, no_labels=pair.no_labels, **kw)
    g = fr.Solver(p)
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, p.rho_0)
    return g, o, pair, rng


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _assert_state_equal(gs, os_, what=""):
    """Every field equal bit for bit (raw bits: -0 differs from +0, a NaN equals only itself)."""
    for k in ("v", "rho", "psi", "psi_inv"):
        a, b = _bits(gs[k]), _bits(os_[k])
        if not np.array_equal(a, b):
            bad = np.flatnonzero((a != b).reshape(len(a), -1).any(-1) if a.ndim > 1 else (a != b))
            raise AssertionError(f"{what} field {k}: {len(bad)} voxels differ, first {bad[:5]}: "
                                 f"gpu {a[bad[:3]]} oracle {b[bad[:3]]}")


@pytest.mark.parametrize("dims,seed", [((24, 20, 18), 3), ((67, 9, 11), 4), ((12, 31, 8), 5),
                                       # the smallest volumes and single-voxel-thick interiors, more
                                       # than two x tiles, more z planes than one chunk
                                       ((3, 3, 3), 11), ((3, 17, 4), 12), ((131, 3, 5), 13), ((4, 5, 66), 14)])
def test_step_teacher_forced_bit_exact(oracle_mod, dims, seed):
    """From an identical random (smooth, large-amplitude) state, one iteration on the GPU equals
    one oracle iteration in every field and in the max update norm."""
    say("step", dims, seed)
    O = oracle_mod
    g, o, pair, rng = _both(dims, O, seed)
    st = phantoms.random_state(dims, rng, amp=1.5)
    g.set_state(**st)
    o.set_state(**st)
    for k, t in enumerate([0.004, 0.008, 0.5]):
        mg, ag = g.step(t)
        mo, ao = o.step(t)
        _assert_state_equal(g.get_state(), o.get_state(), f"step {k}")
        assert mg == mo and ag == ao, (mg, mo, ag, ao)


def test_step_with_signed_zeros_and_underflowing_velocities(oracle_mod):
    """A state whose velocity components include -0, +0 and values whose squares underflow
    (+-1e-30), next to ordinary ones: the (Dv)v diagonal terms must give +0 where the reference's
    fmaxf/fminf form does, so every field matches the oracle in its raw bits, zero signs included."""
    O = oracle_mod
    dims = (40, 22, 18)
    say("signed zeros", dims)
    g, o, pair, rng = _both(dims, O, 33)
    st = phantoms.random_state(dims, rng, amp=1.5)
    v = st["v"].copy()
    comp = v[:, :3]
    pick = rng.random(comp.shape)
    comp[pick < 0.2] = -0.0
    comp[(pick >= 0.2) & (pick < 0.3)] = 0.0
    comp[(pick >= 0.3) & (pick < 0.4)] = -1e-30
    comp[(pick >= 0.4) & (pick < 0.5)] = 1e-30
    st["v"] = v
    g.set_state(**st)
    o.set_state(**st)
    _assert_state_equal(g.get_state(), o.get_state(), "set_state")
    for k, t in enumerate([0.004, 0.5]):
        mg, ag = g.step(t)
        mo, ao = o.step(t)
        _assert_state_equal(g.get_state(), o.get_state(), f"step {k}")
        assert mg == mo and ag == ao, (mg, mo, ag, ao)


@pytest.mark.parametrize("dims,seed,amp", [((70, 40, 36), 31, 4.0), ((131, 19, 25), 32, 9.0)])
def test_step_with_large_displacements_bit_exact(oracle_mod, dims, seed, amp):
    """States displaced by several voxels: most voxels' psi_inv and psi leave the LDS stages, so
    pass A's and pass B's gather fallbacks carry the step; bit-exact against the oracle."""
    say("large-displacement step", dims, seed, amp)
    O = oracle_mod
    g, o, pair, rng = _both(dims, O, seed)
    st = phantoms.random_state(dims, rng, amp=amp, smooth=3.0)
    g.set_state(**st)
    o.set_state(**st)
    for k, t in enumerate([0.004, 0.3]):
        mg, ag = g.step(t)
        mo, ao = o.step(t)
        _assert_state_equal(g.get_state(), o.get_state(), f"step {k}")
        assert mg == mo and ag == ao, (mg, mo, ag, ao)


def _edge_taps(dims, psi_inv):
    """Interior voxels of columns x = 1 and x = X - 2 whose psi_inv takes its trilinear corner on
    the boundary column (x = 0) or past the last interior one (x >= X - 2)."""
    X, Y, Z = dims
    px = psi_inv[:, 0].reshape(Z, Y, X)[1:-1, 1:-1, :]
    return int((np.floor(px[..., 1]) == 0).sum()), int((np.floor(px[..., X - 2]) >= X - 2).sum())


@pytest.mark.parametrize("dims,seed,amp,smooth", [((64, 21, 13), 41, 0.6, 2.0), ((63, 17, 19), 42, 0.9, 2.0),
                                                  ((64, 12, 10), 43, 4.0, 3.0)])
def test_step_one_tile_rows_bit_exact(oracle_mod, dims, seed, amp, smooth):
    """Volumes 63 and 64 voxels wide run pass A as one x tile per row (lane l at x = l, the edge lanes
    on the boundary columns).  Teacher-forced steps from states whose psi_inv takes the trilinear
    taps of voxels next to the boundary onto the edge lanes (and, at the larger amplitude, out of
    the stage into the recovery) equal the oracle bit for bit."""
    say("one-tile step", dims, seed, amp)
    O = oracle_mod
    g, o, pair, rng = _both(dims, O, seed)
    st = phantoms.random_state(dims, rng, amp=amp, smooth=smooth)
    lo, hi = _edge_taps(dims, st["psi_inv"])
    assert lo > 0 and hi > 0, (lo, hi)
    g.set_state(**st)
    o.set_state(**st)
    for k, t in enumerate([0.004, 0.3]):
        mg, ag = g.step(t)
        mo, ao = o.step(t)
        _assert_state_equal(g.get_state(), o.get_state(), f"step {k}")
        assert mg == mo and ag == ao, (mg, mo, ag, ao)


def test_run_bit_exact_with_log(oracle_mod):
    """A 25-iteration verbose run: state, max-norm history, log lines and Dice equal the oracle."""
    O = oracle_mod
    dims = (32, 28, 24)
    say("run", dims)
    g, o, pair, rng = _both(dims, O, 7, max_iter=25, verbosity=2, alpha=0.4)
    g.run()
    prm = O.SolverParams(max_iter=25, verbosity=2, no_labels=pair.no_labels, alpha=0.4)
    r = o.run(prm)
    st = g.stats()
    assert st.iterations == r["iterations"] == 25
    mh, eh = g.history()
    assert np.array_equal(mh, r["max_hist"])
    np.testing.assert_allclose(eh, r["energy_hist"][:25], rtol=1e-12)
    _assert_state_equal(g.get_state(), o.get_state(), "run")
    outs = g.finalize()
    oo = o.outputs()
    for a, b, nm in zip(outs, oo, ("img1", "lab1", "img2", "lab2")):
        assert np.array_equal(a, b), nm
    dice, ov, vx = g.dice()
    oov, ovx, _ = o.dice_counts(pair.no_labels)
    assert np.array_equal(ov, oov) and np.array_equal(vx, ovx)
    mt, mr = o.min_det()
    st = g.stats()
    assert st.min_det_true == mt and st.min_det_ref == mr


@pytest.mark.parametrize("dims,seed", [((3, 3, 3), 21), ((5, 4, 3), 22), ((130, 11, 4), 23)])
def test_degenerate_volume_run_bit_exact(oracle_mod, dims, seed):
    """Verbose 12-iteration runs on volumes with one-voxel-thick interiors: state, max-norm
    history (argmax 0 when every update is 0), energies, outputs and Dice equal the oracle."""
    O = oracle_mod
    say("degenerate run", dims, seed)
    g, o, pair, rng = _both(dims, O, seed, max_iter=12, verbosity=2, alpha=0.4)
    g.run()
    r = o.run(O.SolverParams(max_iter=12, verbosity=2, no_labels=pair.no_labels, alpha=0.4))
    st = g.stats()
    assert st.iterations == r["iterations"]
    mh, eh = g.history()
    assert np.array_equal(mh, r["max_hist"])
    np.testing.assert_allclose(eh, r["energy_hist"][:len(eh)], rtol=1e-12)
    _assert_state_equal(g.get_state(), o.get_state(), "run")
    for a, b, nm in zip(g.finalize(), o.outputs(), ("img1", "lab1", "img2", "lab2")):
        assert np.array_equal(a, b), nm
    dice, ov, vx = g.dice()
    oov, ovx, _ = o.dice_counts(pair.no_labels)
    assert np.array_equal(ov, oov) and np.array_equal(vx, ovx)


@pytest.mark.parametrize("no_labels,dims", [(9, (70, 21, 12)), (3000, (130, 9, 7))])
def test_dice_counts_with_labels_out_of_range(oracle_mod, no_labels, dims):
    """Dice counts (per label: voxels of each volume and agreeing voxels) over labels that include
    values outside [0, no_labels), with the counts held in LDS (9 labels) and in global memory
    (3000 labels): equal to the oracle's."""
    O = oracle_mod
    X, Y, Z = dims
    say("dice counts", no_labels, dims)
    rng = np.random.default_rng(no_labels)
    base = phantoms.blobs(shape=(Z, Y, X), seed=51)
    # runs of equal labels (as in real label maps) with scattered out-of-range values
    lab = np.repeat(rng.integers(-2, no_labels + 2, (Z * Y * X + 6) // 7), 7)[:Z * Y * X]
    lab2 = lab.copy()
    flip = rng.random(lab.size) < 0.2
    lab2[flip] = rng.integers(-1, no_labels + 1, int(flip.sum()))
    pair = phantoms.Pair(base.target, lab.astype(np.int16).reshape(Z, Y, X), base.source,
                         lab2.astype(np.int16).reshape(Z, Y, X), no_labels)
    g, o, _, _ = _both(dims, O, 51, src_pair=pair, max_iter=2, alpha=0.3)
    g.run()
    g.finalize()
    o.run(O.SolverParams(max_iter=2, alpha=0.3, no_labels=no_labels))
    o.outputs()
    _, ov, vx = g.dice()
    oov, ovx, obad = o.dice_counts(no_labels)
    assert obad > 0
    assert np.array_equal(ov, oov) and np.array_equal(vx, ovx)


@pytest.mark.parametrize("keep", [False, True])
def test_second_registration_on_one_context(oracle_mod, keep):
    """Two registrations in a row on one context, the second of another pair.  With
    fr_set_keep_state the second continues from the v, psi and psi_inv the first left (new images,
    rho = rho_0), as the reference's FastReg does from one call to the next; without it the second
    starts afresh.  Outputs and state bit-exact against the oracle doing the same."""
    O = oracle_mod
    dims = (30, 26, 22)
    X, Y, Z = dims
    say("second registration", keep)
    a, b = phantoms.blobs(shape=(Z, Y, X), seed=61), phantoms.blobs(shape=(Z, Y, X), seed=62)
    nl = max(a.no_labels, b.no_labels)
    p = _params(dims, no_labels=nl, max_iter=15, alpha=0.3)
    g = fr.Solver(p)
    g.set_keep_state(keep)
    assert g.keep_state() == keep
    g.register(a.target, a.target_labels, a.source, a.source_labels)
    outs = g.register(b.target, b.target_labels, b.source, b.source_labels)
    o = O.Oracle(dims, "f32")
    prm = O.SolverParams(max_iter=15, alpha=0.3, no_labels=nl)
    o.init(a.target, a.target_labels, a.source, a.source_labels)
    o.run(prm)
    st = o.get_state()
    o.init(b.target, b.target_labels, b.source, b.source_labels)
    if keep:
        o.set_state(v=st["v"], psi=st["psi"], psi_inv=st["psi_inv"])
    o.run(prm)
    for x, y, nm in zip(outs, o.outputs(), ("img1", "lab1", "img2", "lab2")):
        assert np.array_equal(np.asarray(x).reshape(-1), y), nm
    _assert_state_equal(g.get_state(), o.get_state(), "second registration")


def test_register_matches_oracle_c1_shape(oracle_mod):
    """fr_register on the 64^3 blob phantom (40 iterations): in-place outputs, Dice and
    min det J equal the oracle's."""
    O = oracle_mod
    pair = phantoms.blobs()
    dims = pair.dims
    say("register", dims)
    p = _params(dims, no_labels=pair.no_labels, max_iter=40)
    g = fr.Solver(p)
    outs = g.register(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o.run(O.SolverParams(max_iter=40, no_labels=pair.no_labels))
    for a, b, nm in zip(outs, o.outputs(), ("img1", "lab1", "img2", "lab2")):
        assert np.array_equal(a.reshape(-1), b), nm
    _, ov, vx = g.dice()
    oov, ovx, _ = o.dice_counts(pair.no_labels)
    assert np.array_equal(ov, oov) and np.array_equal(vx, ovx)
    st = g.stats()
    assert (st.min_det_true, st.min_det_ref) == o.min_det()
    assert st.iterations == 40 and st.stop_reason == 0


def test_energy_matches_oracle(oracle_mod):
    O = oracle_mod
    dims = (30, 26, 22)
    g, o, pair, rng = _both(dims, O, 9)
    st = phantoms.random_state(dims, rng, amp=2.0)
    g.set_state(**st)
    o.set_state(**st)
    ge = g.energy()
    oe = o.energy()
    np.testing.assert_allclose(ge, oe, rtol=1e-12)


def test_stop_rules(oracle_mod):
    O = oracle_mod
    dims = (20, 18, 16)
    # converged after the first iteration (MAX_UPDATE_NORM huge)
    g, o, pair, _ = _both(dims, O, 11, max_iter=50, max_update_norm=1e30)
    g.run()
    st = g.stats()
    r = o.run(O.SolverParams(max_iter=50, max_update_norm=1e30, no_labels=pair.no_labels))
    assert st.iterations == r["iterations"] == 1 and st.stop_reason == r["stop_reason"] == 1
    assert "SOLVER CONVERGED AFTER 1 ITERATIONS" in g.log()
    _assert_state_equal(g.get_state(), o.get_state(), "converged")
    # RESET_ITER t schedule
    g, o, pair, _ = _both(dims, O, 12, max_iter=9, reset_iter=3, alpha=0.3)
    g.run()
    o.run(O.SolverParams(max_iter=9, reset_iter=3, alpha=0.3, no_labels=pair.no_labels))
    _assert_state_equal(g.get_state(), o.get_state(), "reset_iter")
    # NaN energy stops before the first update (verbosity 1)
    X, Y, Z = dims
    tgt = pair.target.copy()
    tgt.reshape(-1)[5] = np.nan
    p = _params(dims, no_labels=pair.no_labels, max_iter=5, verbosity=1)
    g = fr.Solver(p)
    g.init(tgt, pair.target_labels, pair.source, pair.source_labels)
    g.run()
    st = g.stats()
    assert st.iterations == 0 and st.stop_reason == 2
    o = O.Oracle(dims, "f32")
    o.init(tgt, pair.target_labels, pair.source, pair.source_labels)
    r = o.run(O.SolverParams(max_iter=5, verbosity=1, no_labels=pair.no_labels))
    assert r["iterations"] == 0 and r["stop_reason"] == 2


def test_full_size_brain_three_iterations(oracle_mod):
    """At the full 182x218x182 size: three GPU iterations equal three oracle iterations."""
    O = oracle_mod
    pair = phantoms.brain()
    dims = pair.dims
    say("brain", dims)
    p = _params(dims, no_labels=pair.no_labels, max_iter=3, alpha=0.5)
    g = fr.Solver(p)
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    g.run()
    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    r = o.run(O.SolverParams(max_iter=3, alpha=0.5, no_labels=pair.no_labels))
    mh, _ = g.history()
    assert np.array_equal(mh, r["max_hist"])
    _assert_state_equal(g.get_state(), o.get_state(), "brain")


@pytest.mark.parametrize("zc", [(1, 1), (5, 2), (0, 0), (7, 40), (60, 16), (2, 3)])
def test_chunking_does_not_change_results(oracle_mod, zc):
    """Every z-chunking of the two iteration kernels gives the oracle's iteration bit for bit
    (chunk edges fall on different slices for each choice, including a single chunk and chunks
    of one slice)."""
    O = oracle_mod
    dims = (70, 21, 40)
    say("chunking", zc)
    g, o, pair, rng = _both(dims, O, 13)
    g.set_chunking(*zc)
    got = g.chunking()
    assert all(got[k] == zc[k] for k in range(2) if zc[k]), got
    st = phantoms.random_state(dims, rng, amp=1.5)
    g.set_state(**st)
    o.set_state(**st)
    for t in (0.01, 0.3):
        mg, ag = g.step(t)
        mo, ao = o.step(t)
        assert mg == mo and ag == ao, (mg, mo, ag, ao)
    _assert_state_equal(g.get_state(), o.get_state(), f"zc {zc}")


def test_convergence_at_an_exact_threshold(oracle_mod):
    """MAX_UPDATE_NORM set to the oracle's max norm of iteration 5: the GPU stops at the same
    iteration with the same message.  The device-side test must read the maximum over all of the
    iteration's slots (a partial maximum would stop earlier)."""
    O = oracle_mod
    dims = (70, 60, 50)
    say("threshold", dims)
    g0, o0, pair, _ = _both(dims, O, 21, max_iter=12, alpha=0.3)
    mh = o0.run(O.SolverParams(max_iter=12, alpha=0.3, no_labels=pair.no_labels))["max_hist"]
    thr = float(mh[4])
    g, o, pair, _ = _both(dims, O, 21, max_iter=12, alpha=0.3, max_update_norm=thr)
    r = o.run(O.SolverParams(max_iter=12, alpha=0.3, no_labels=pair.no_labels, max_update_norm=thr))
    g.run()
    st = g.stats()
    assert r["stop_reason"] == 1 and r["iterations"] <= 5
    assert st.iterations == r["iterations"] and st.stop_reason == r["stop_reason"], (st.iterations, r)
    assert f"SOLVER CONVERGED AFTER {r['iterations']} ITERATIONS" in g.log()
    _assert_state_equal(g.get_state(), o.get_state(), "threshold")
