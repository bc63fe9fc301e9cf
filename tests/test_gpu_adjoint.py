"""The differentiated solver (fr_adjoint_*) on the GPU: its FP64 forward pass bit-exact against the
FP64 oracle iteration by iteration, its gradient (hand-written adjoint kernels, checkpointed)
against torch autograd of the unrolled restatement (oracle/unrolled.py) and against central
finite differences of its own forward pass, independent of the checkpoint interval."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

import fastreg_amd as fr
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu

ALPHA = float(np.float32(0.004))
TAU = float(np.float32(0.2))
W_REG = float(np.float32(0.2))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(dims, seed):
    X, Y, Z = dims
    return phantoms.blobs(shape=(Z, Y, X), seed=seed, n_blobs=5, n_labels=4, n_rbf=3, rbf_sigma=4.0, max_v=1.5,
                          steps=4)


def _v0(dims, seed, amp=1.0, off=0.1):
    """A smooth velocity, nonzero on the boundary too (v = 0 would sit on every upwind kink)."""
    return phantoms.smooth_velocity(dims, seed, amp, off)


def _adjoint(dims, pair, K, ckpt, **kw):
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, alpha=ALPHA, **kw)
    a = fr.Adjoint(p, K, ckpt)
    a.set_images(pair.target, pair.target_labels, pair.source, pair.source_labels)
    return a


@pytest.mark.parametrize("dims,seed,K,reset_iter", [((24, 20, 18), 3, 5, -1), ((19, 13, 11), 4, 6, 3)])
def test_forward_is_bit_exact_against_the_fp64_oracle(oracle_mod, dims, seed, K, reset_iter):
    O = oracle_mod
    say("adjoint forward", dims, K, reset_iter)
    n = dims[0] * dims[1] * dims[2]
    pair = _pair(dims, seed)
    v0 = _v0(dims, seed)
    a = _adjoint(dims, pair, K, 1, reset_iter=reset_iter)
    try:
        a.forward(v0, W_REG, TAU, 0.0)
        o = O.Oracle(dims, "f64")
        o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, 1.0)
        v4 = np.zeros((n, 4))
        v4[:, :3] = v0
        o.set_state(v=v4)
        g0 = a.state(0)
        assert np.array_equal(g0["v"], v0) and np.all(g0["rho"] == 1.0)
        t = ALPHA
        for k in range(1, K + 1):
            if k >= 2:
                t += ALPHA
                if reset_iter > 0 and k % reset_iter == 0:
                    t = ALPHA
            o.step(t, ALPHA, TAU, W_REG, 2.0, 0.25)
            ref, got = o.get_state(), a.state(k)
            for key in ("v", "psi", "psi_inv"):
                assert np.array_equal(got[key], ref[key][:, :3]), (k, key)
            assert np.array_equal(got["rho"], ref["rho"]), (k, "rho")
    finally:
        a.close()


def test_forward_at_the_workload_size_is_bit_exact_against_the_fp64_oracle(oracle_mod):
    """The differentiated workload's own forward pass: the 128^3 blob pair, K = 64, checkpoints
    every 8, from its smooth v0; every kept state (iterations 8, 16, .., 64) equals the FP64
    oracle's, run on the host the same way.  With the gradient equal to autograd of the unrolled
    restatement at this size (the full-size test below), the gradient is that of the oracle's
    function."""
    O = oracle_mod
    dims, K, ckpt = (128, 128, 128), 64, 8
    X, Y, Z = dims
    n = X * Y * Z
    say("adjoint forward vs FP64 oracle, full size", dims, K)
    pair = phantoms.blobs(shape=(Z, Y, X), seed=5, scale=2.0)
    v0 = phantoms.smooth_velocity(dims, 5, amp=0.5, offset=0.05, sigma=4.0)
    a = _adjoint(dims, pair, K, ckpt)
    o = O.Oracle(dims, "f64")
    try:
        a.forward(v0, W_REG, TAU, 1.0)
        o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, 1.0)
        v4 = np.zeros((n, 4))
        v4[:, :3] = v0
        o.set_state(v=v4)
        t = ALPHA
        for k in range(1, K + 1):
            if k >= 2:
                t += ALPHA
            o.step(t, ALPHA, TAU, W_REG, 2.0, 0.25)
            if k % ckpt == 0:
                ref, got = o.get_state(), a.state(k)
                for key in ("v", "psi", "psi_inv"):
                    assert np.array_equal(got[key], ref[key][:, :3]), (k, key)
                assert np.array_equal(got["rho"], ref["rho"]), (k, "rho")
                say("  iteration", k, "equal")
    finally:
        o.close()
        a.close()


def test_gradient_matches_autograd_of_the_unrolled_restatement():
    from oracle import unrolled as U
    dims, K, ckpt, lw = (20, 18, 16), 7, 3, 0.5
    say("adjoint vs autograd", dims, K, ckpt)
    pair = _pair(dims, 5)
    v0 = _v0(dims, 5)
    a = _adjoint(dims, pair, K, ckpt)
    try:
        e = a.forward(v0, W_REG, TAU, lw)
        g, gw, gt = a.backward()
    finally:
        a.close()
    v0t = torch.tensor(v0, requires_grad=True)
    wt = torch.tensor(W_REG, dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    E = U.unrolled(dims, pair.target, pair.target_labels, pair.source, pair.source_labels, v0t, wt, tt, K,
                   alpha=ALPHA, no_labels=pair.no_labels, label_weight=lw)
    E.backward()
    assert abs(e - E.item()) <= 1e-12 * abs(E.item())
    gr = v0t.grad.numpy()
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    say("rel L2 of dE/dv0", rel, "w_reg", gw, wt.grad.item(), "tau", gt, tt.grad.item())
    assert rel <= 1e-10
    assert abs(gw - wt.grad.item()) <= 1e-10 * abs(wt.grad.item())
    assert abs(gt - tt.grad.item()) <= 1e-10 * abs(tt.grad.item())


def _autograd(dims, pair, v0, K, lw, device="cpu", checkpoint=False):
    from oracle import unrolled as U
    v0t = torch.tensor(v0, device=device, requires_grad=True)
    wt = torch.tensor(W_REG, dtype=torch.float64, device=device, requires_grad=True)
    tt = torch.tensor(TAU, dtype=torch.float64, device=device, requires_grad=True)
    E = U.unrolled(dims, pair.target, pair.target_labels, pair.source, pair.source_labels, v0t, wt, tt, K,
                   alpha=ALPHA, no_labels=pair.no_labels, label_weight=lw, checkpoint=checkpoint)
    E.backward()
    return E.item(), v0t.grad.cpu().numpy(), wt.grad.item(), tt.grad.item()


def _compare(res, ref, tol):
    e, g, gw, gt = res
    E, G, GW, GT = ref
    rel = np.linalg.norm(g - G) / np.linalg.norm(G)
    say("E", e, E, "rel L2 dE/dv0", rel, "w_reg", gw, GW, "tau", gt, GT)
    assert abs(e - E) <= 1e-12 * abs(E)
    assert rel <= tol
    assert abs(gw - GW) <= tol * abs(GW) and abs(gt - GT) <= tol * abs(GT)


def test_gradient_matches_autograd_at_depth_64():
    """K = 64 unrolled iterations with checkpoints every 8 (the workload's depth) on a small volume."""
    dims, K, ckpt, lw = (24, 20, 18), 64, 8, 1.0
    say("adjoint vs autograd, depth", dims, K, ckpt)
    pair = _pair(dims, 12)
    v0 = _v0(dims, 12, amp=0.5, off=0.05)
    a = _adjoint(dims, pair, K, ckpt)
    try:
        e = a.forward(v0, W_REG, TAU, lw)
        g, gw, gt = a.backward()
    finally:
        a.close()
    _compare((e, g, gw, gt), _autograd(dims, pair, v0, K, lw), 1e-9)


def test_gradient_with_uneven_gather_tiles():
    """Shapes whose record gathers (64 x 4 tiles marching z-chunks) end in partial tiles and a
    partial last z-chunk, whatever the GPU's CU count: 70 x 66 x 71 with 5-plane chunks (2 x 17
    tiles, the last x-tile 6 columns wide, the last y-tile 2 rows high, the last chunk one plane);
    130 x 7 x 9 with 4-plane chunks (a 2-column last x-tile, a 3-row last y-tile, a one-plane last
    chunk)."""
    for dims, K, ckpt, zc in (((70, 66, 71), 4, 2, 5), ((130, 7, 9), 5, 5, 4)):
        say("adjoint vs autograd, uneven gather tiles", dims, K, ckpt, zc)
        pair = _pair(dims, 14)
        v0 = _v0(dims, 14)
        a = _adjoint(dims, pair, K, ckpt)
        try:
            a.set_gather_chunk(zc)
            tx, ty, got = a.gather_chunk()
            assert got == zc
            assert dims[0] % tx != 0 and dims[1] % ty != 0 and dims[2] % zc == 1, (tx, ty, zc)
            e = a.forward(v0, W_REG, TAU, 0.5)
            g, gw, gt = a.backward()
        finally:
            a.close()
        _compare((e, g, gw, gt), _autograd(dims, pair, v0, K, 0.5), 1e-10)


def test_scalar_gradients_are_the_same_bits_on_every_run():
    """dE/dW_REG and dE/dTAU are sums over blocks added in a fixed order (no atomics): two backward
    passes of the same forward pass return the same bits, and so does the whole gradient when no
    source is displaced by a voxel or more (no atomic scatter runs)."""
    dims, K, ckpt = (40, 36, 30), 6, 3
    say("adjoint determinism", dims, K)
    pair = _pair(dims, 15)
    v0 = _v0(dims, 15, amp=0.5, off=0.05)
    a = _adjoint(dims, pair, K, ckpt)
    try:
        a.forward(v0, W_REG, TAU, 1.0)
        g1, gw1, gt1 = a.backward()
        a.forward(v0, W_REG, TAU, 1.0)
        g2, gw2, gt2 = a.backward()
    finally:
        a.close()
    assert gw1 == gw2 and gt1 == gt2
    assert np.array_equal(g1, g2)


def test_step_pairs_reject_overlapping_arrays():
    """fr_step_fwd's outputs may not overlap its inputs or each other, and fr_step_bwd's adjoint
    arrays may not overlap each other: both return FR_ERR_INVALID before running anything."""
    dims = (12, 10, 9)
    n = dims[0] * dims[1] * dims[2]
    pair = _pair(dims, 16)
    a = _adjoint(dims, pair, 2, 1)
    idx = np.arange(n)
    ident = np.stack([idx % dims[0], (idx // dims[0]) % dims[1], idx // (dims[0] * dims[1])], 1).astype(np.float64)
    state = {"rho": np.ones(n), "v": _v0(dims, 16), "psi": ident, "psi_inv": ident}
    bufs = a._state_bufs(state) + [fr.DeviceArray.from_host(np.zeros(w * n)) for w in (1, 3, 3, 3)]
    L = fr.lib()
    try:
        ok = [b.ptr for b in bufs]
        assert L.fr_step_fwd(a.h, 1, W_REG, TAU, *ok) == 0
        for o, i in ((5, 1), (4, 0), (7, 3), (6, 2)):
            bad = list(ok)
            bad[o] = ok[i]          # an output in place of its own input
            assert L.fr_step_fwd(a.h, 1, W_REG, TAU, *bad) == fr.FR_ERR_INVALID
            assert b"overlaps" in L.fr_last_error()
        bad = list(ok)
        bad[7] = ok[6]              # two outputs on the same memory
        assert L.fr_step_fwd(a.h, 1, W_REG, TAU, *bad) == fr.FR_ERR_INVALID
        gw, gt = C.c_double(), C.c_double()
        assert L.fr_step_bwd(a.h, 1, W_REG, TAU, *ok, C.byref(gw), C.byref(gt)) == 0
        bad = list(ok)
        bad[6] = ok[5]              # two adjoint arrays on the same memory
        assert L.fr_step_bwd(a.h, 1, W_REG, TAU, *bad, C.byref(gw), C.byref(gt)) == fr.FR_ERR_INVALID
    finally:
        for b in bufs:
            b.free()
        a.close()


def test_gradient_with_displacements_of_several_voxels():
    """A large v0 moves psi and psi_inv by more than a voxel from the first iteration on, so that the
    transposes of the trilinear samples take their atomic path for the displaced sources."""
    dims, K, lw = (20, 18, 16), 3, 0.5
    say("adjoint vs autograd, large displacements", dims, K)
    n = dims[0] * dims[1] * dims[2]
    pair = _pair(dims, 13)
    v0 = phantoms.smooth_velocity(dims, 13, amp=200.0, offset=20.0)
    a = _adjoint(dims, pair, K, 1)
    try:
        e = a.forward(v0, W_REG, TAU, lw)
        g, gw, gt = a.backward()
        idx = np.arange(n)
        ident = np.stack([idx % dims[0], (idx // dims[0]) % dims[1], idx // (dims[0] * dims[1])], 1)
        for k in (1, 2):
            st = a.state(k)
            say("state", k, "max displacement psi", np.abs(st["psi"] - ident).max(),
                "psi_inv", np.abs(st["psi_inv"] - ident).max())
            assert np.abs(st["psi"] - ident).max() > 1.0 and np.abs(st["psi_inv"] - ident).max() > 1.0
    finally:
        a.close()
    _compare((e, g, gw, gt), _autograd(dims, pair, v0, K, lw), 1e-9)


def test_full_size_gradient_matches_autograd_on_the_gpu():
    """The workload itself: the 128^3 blob pair, K = 64, checkpoints every 8, against torch autograd
    of the unrolled restatement run on the GPU in FP64 (each iteration recomputed in its backward
    pass, torch.utils.checkpoint, to bound memory)."""
    if not torch.cuda.is_available():
        pytest.fail("torch sees no GPU")
    dims, K, ckpt, lw = (128, 128, 128), 64, 8, 1.0
    say("adjoint vs autograd, full size", dims, K, ckpt)
    X, Y, Z = dims
    pair = phantoms.blobs(shape=(Z, Y, X), seed=5, scale=2.0)
    v0 = phantoms.smooth_velocity(dims, 5, amp=0.5, offset=0.05, sigma=4.0)
    a = _adjoint(dims, pair, K, ckpt)
    try:
        e = a.forward(v0, W_REG, TAU, lw)
        g, gw, gt = a.backward()
    finally:
        a.close()
    say("adjoint done")
    ref = _autograd(dims, pair, v0, K, lw, device="cuda", checkpoint=True)
    torch.cuda.synchronize()
    _compare((e, g, gw, gt), ref, 1e-9)


def test_gradient_matches_central_finite_differences():
    """Central differences of the library's own forward pass.  The solver is piecewise smooth: its
    upwind schemes switch branch where a velocity component changes sign, and the trilinear warps
    are piecewise linear between grid nodes, which psi (starting at the identity) stays close to.
    Steps of 1e-6 and more already cross such kinks at this size, so the steps are 1e-7 (v0) and
    1e-6 (W_REG, TAU); the tolerance includes the rounding of E over the step."""
    dims, K, ckpt, lw = (28, 24, 20), 10, 4, 0.5
    say("adjoint vs FD", dims, K, ckpt)
    n = dims[0] * dims[1] * dims[2]
    pair = _pair(dims, 6)
    v0 = _v0(dims, 6)
    a = _adjoint(dims, pair, K, ckpt)
    try:
        e = a.forward(v0, W_REG, TAU, lw)
        g, gw, gt = a.backward()
        rng = np.random.default_rng(8)

        def close(fd, an, eps):
            return abs(fd - an) <= 1e-4 * abs(an) + 8 * abs(e) * 2.2e-16 / eps

        eps = 1e-7
        for _ in range(4):
            d = rng.normal(size=(n, 3))
            fd = (a.forward(v0 + eps * d, W_REG, TAU, lw) - a.forward(v0 - eps * d, W_REG, TAU, lw)) / (2 * eps)
            an = float((g * d).sum())
            say("direction", fd, an)
            assert close(fd, an, eps), (fd, an)
        eps = 1e-6
        fd = (a.forward(v0, W_REG + eps, TAU, lw) - a.forward(v0, W_REG - eps, TAU, lw)) / (2 * eps)
        say("w_reg", fd, gw)
        assert close(fd, gw, eps), (fd, gw)
        fd = (a.forward(v0, W_REG, TAU + eps, lw) - a.forward(v0, W_REG, TAU - eps, lw)) / (2 * eps)
        say("tau", fd, gt)
        assert close(fd, gt, eps), (fd, gt)
    finally:
        a.close()


def test_gradient_matches_finite_differences_at_eps_1e4_on_the_ramp_pair():
    """Central differences at the full step of 1e-4 where no branch switches: on the ramp pair,
    from a small positive v0, every velocity component stays positive, psi stays ahead of and
    psi_inv behind the grid nodes, so no upwind branch, trilinear cell or clamp changes between
    the two evaluations, and FD agrees with the adjoint to 1e-6 (plus the rounding of E over the
    step)."""
    dims, K, ckpt, lw = (40, 36, 32), 32, 8, 1.0
    say("adjoint vs FD on the ramp pair", dims, K, ckpt)
    X, Y, Z = dims
    n = X * Y * Z
    pair = phantoms.ramp((Z, Y, X))
    v0 = phantoms.ramp_velocity(dims, 3)
    a = _adjoint(dims, pair, K, ckpt)
    try:
        e = a.forward(v0, W_REG, TAU, lw)
        g, gw, gt = a.backward()
        eps = 1e-4

        def close(fd, an):
            return abs(fd - an) <= 1e-6 * abs(an) + 8 * abs(e) * 2.2e-16 / eps

        rng = np.random.default_rng(10)
        for _ in range(4):
            d = rng.normal(size=(n, 3))
            fd = (a.forward(v0 + eps * d, W_REG, TAU, lw) - a.forward(v0 - eps * d, W_REG, TAU, lw)) / (2 * eps)
            an = float((g * d).sum())
            say("direction", fd, an)
            assert close(fd, an), (fd, an)
        fd = (a.forward(v0, W_REG + eps, TAU, lw) - a.forward(v0, W_REG - eps, TAU, lw)) / (2 * eps)
        say("w_reg", fd, gw)
        assert close(fd, gw), (fd, gw)
        fd = (a.forward(v0, W_REG, TAU + eps, lw) - a.forward(v0, W_REG, TAU - eps, lw)) / (2 * eps)
        say("tau", fd, gt)
        assert close(fd, gt), (fd, gt)
    finally:
        a.close()


def test_checkpoint_interval_does_not_change_the_gradient():
    dims, K, lw = (22, 17, 15), 9, 0.25
    say("adjoint checkpoint intervals", dims, K)
    pair = _pair(dims, 9)
    v0 = _v0(dims, 9)
    res = {}
    for ckpt in (1, 4, 9):
        a = _adjoint(dims, pair, K, ckpt)
        try:
            e = a.forward(v0, W_REG, TAU, lw)
            g, gw, gt = a.backward()
            st = a.state(K)
            res[ckpt] = (e, g, gw, gt, st)
        finally:
            a.close()
    e1, g1, gw1, gt1, st1 = res[1]
    for ckpt in (4, 9):
        e, g, gw, gt, st = res[ckpt]
        assert e == e1
        for key in st1:
            assert np.array_equal(st[key], st1[key])
        # the adjoint's scatters are FP64 atomics, whose order varies
        assert np.linalg.norm(g - g1) <= 1e-12 * np.linalg.norm(g1)
        assert abs(gw - gw1) <= 1e-12 * abs(gw1) and abs(gt - gt1) <= 1e-12 * abs(gt1)


def test_autograd_function_wrapper_and_repeat_calls():
    from adjoint_function import SolverObjective
    dims, K, ckpt, lw = (12, 10, 9), 4, 2, 0.5
    say("adjoint autograd.Function", dims)
    pair = _pair(dims, 10)
    v0 = _v0(dims, 10)
    a = _adjoint(dims, pair, K, ckpt)
    try:
        e_direct = a.forward(v0, W_REG, TAU, lw)
        g_direct, gw_direct, gt_direct = a.backward()
        v = torch.tensor(v0, requires_grad=True)
        w = torch.tensor(W_REG, dtype=torch.float64, requires_grad=True)
        t = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
        E = SolverObjective.apply(v, w, t, a, lw)
        assert E.item() == e_direct
        (2.0 * E).backward()
        assert np.linalg.norm(v.grad.numpy() - 2 * g_direct) <= 1e-12 * np.linalg.norm(g_direct)
        assert abs(w.grad.item() - 2 * gw_direct) <= 1e-12 * abs(gw_direct)
        assert abs(t.grad.item() - 2 * gt_direct) <= 1e-12 * abs(gt_direct)
        ok = torch.autograd.gradcheck(lambda vv: SolverObjective.apply(vv, w.detach(), t.detach(), a, lw),
                                      (torch.tensor(v0, requires_grad=True),), eps=1e-6, atol=1e-8, rtol=1e-5,
                                      nondet_tol=1e-12, fast_mode=True)
        assert ok
        by, fwd_ms, bwd_ms = a.info()
        assert by > 0 and fwd_ms > 0 and bwd_ms > 0
        with pytest.raises(fr.FastRegError):
            a.state(1)  # not a kept state (ckpt 2)
    finally:
        a.close()


def test_backward_before_forward_is_a_state_error():
    dims = (10, 9, 8)
    pair = _pair(dims, 11)
    a = _adjoint(dims, pair, 3, 2)
    try:
        with pytest.raises(fr.FastRegError) as e:
            a.backward()
        assert e.value.code == fr.FR_ERR_STATE
    finally:
        a.close()
