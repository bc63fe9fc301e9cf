"""The per-operator FP64 forward/backward pairs of the C-ABI (fr_warp_*, fr_gradient_*,
fr_laplacian_*) and the one-iteration pair (fr_step_fwd / fr_step_bwd): forward passes against
the FP64 oracle or a direct restatement, backward passes against torch autograd of the
restatement in oracle/unrolled.py."""
import sys

import numpy as np
import pytest
import torch

import fastreg_amd as fr
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu

ALPHA = float(np.float32(0.004))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def test_warp_pair():
    from oracle import unrolled as U
    dims = (13, 11, 9)
    X, Y, Z = dims
    n = X * Y * Z
    say("warp pair", dims)
    rng = np.random.default_rng(1)
    f = rng.normal(size=n)
    idx = np.arange(n)
    pos = np.stack([idx % X, (idx // X) % Y, idx // (X * Y)], 1) + rng.normal(0, 1.5, (n, 3))
    gout = rng.normal(size=n)
    out = fr.warp_fwd(f, pos, dims)
    g = U._Grid(dims)
    ft = torch.tensor(f, requires_grad=True)
    pt = torch.tensor(pos, requires_grad=True)
    ref = U.trilerp(g, ft, pt)
    np.testing.assert_allclose(out, ref.detach().numpy(), rtol=0, atol=1e-14)
    (ref * torch.tensor(gout)).sum().backward()
    gf, gp = fr.warp_bwd(f, pos, gout, dims)
    assert _rel(gf, ft.grad.numpy()) <= 1e-13
    assert _rel(gp, pt.grad.numpy()) <= 1e-12


def test_gradient_pair():
    dims = (10, 9, 8)
    X, Y, Z = dims
    n = X * Y * Z
    say("gradient pair", dims)
    rng = np.random.default_rng(2)
    f = rng.normal(size=n)
    got = fr.gradient_fwd(f, dims)
    F = f.reshape(Z, Y, X)
    ref = np.zeros((Z, Y, X, 3))
    ref[1:-1, 1:-1, 1:-1, 0] = (F[1:-1, 1:-1, 2:] - F[1:-1, 1:-1, :-2]) / 2.0
    ref[1:-1, 1:-1, 1:-1, 1] = (F[1:-1, 2:, 1:-1] - F[1:-1, :-2, 1:-1]) / 2.0
    ref[1:-1, 1:-1, 1:-1, 2] = (F[2:, 1:-1, 1:-1] - F[:-2, 1:-1, 1:-1]) / 2.0
    assert np.array_equal(got, ref.reshape(n, 3))
    gg = rng.normal(size=(n, 3))
    ft = torch.tensor(F, requires_grad=True)
    out = torch.zeros((Z, Y, X, 3), dtype=torch.float64)
    out[1:-1, 1:-1, 1:-1, 0] = (ft[1:-1, 1:-1, 2:] - ft[1:-1, 1:-1, :-2]) / 2.0
    out[1:-1, 1:-1, 1:-1, 1] = (ft[1:-1, 2:, 1:-1] - ft[1:-1, :-2, 1:-1]) / 2.0
    out[1:-1, 1:-1, 1:-1, 2] = (ft[2:, 1:-1, 1:-1] - ft[:-2, 1:-1, 1:-1]) / 2.0
    (out.reshape(n, 3) * torch.tensor(gg)).sum().backward()
    assert _rel(fr.gradient_bwd(gg, dims), ft.grad.numpy().reshape(n)) <= 1e-14


def test_laplacian_pair(oracle_mod):
    dims = (11, 10, 9)
    X, Y, Z = dims
    n = X * Y * Z
    say("laplacian pair", dims)
    rng = np.random.default_rng(3)
    psi = rng.normal(size=(n, 3))
    got = fr.laplacian_fwd(psi, dims)
    p4 = np.zeros((n, 4))
    p4[:, :3] = psi
    ref = oracle_mod.laplacian(dims, p4, prec="f64")[:, :3]
    assert np.array_equal(got, ref)
    gl = rng.normal(size=(n, 3))
    pt = torch.tensor(psi.reshape(Z, Y, X, 3), requires_grad=True)
    s = pt[1:-1, 1:-1, 1:-1] * -6.0
    for sl in ((slice(1, -1), slice(1, -1), slice(2, None)), (slice(1, -1), slice(1, -1), slice(None, -2)),
               (slice(1, -1), slice(2, None), slice(1, -1)), (slice(1, -1), slice(None, -2), slice(1, -1)),
               (slice(2, None), slice(1, -1), slice(1, -1)), (slice(None, -2), slice(1, -1), slice(1, -1))):
        s = s + pt[sl]
    L = torch.zeros((Z, Y, X, 3), dtype=torch.float64)
    L[1:-1, 1:-1, 1:-1] = s * -1.0
    (L.reshape(n, 3) * torch.tensor(gl)).sum().backward()
    assert _rel(fr.laplacian_bwd(gl, dims), pt.grad.numpy().reshape(n, 3)) <= 1e-14


def _step_case(dims, seed):
    X, Y, Z = dims
    pair = phantoms.blobs(shape=(Z, Y, X), seed=seed, n_blobs=5, n_labels=4, n_rbf=3, rbf_sigma=4.0, max_v=1.5,
                          steps=4)
    rng = np.random.default_rng(seed)
    st = phantoms.random_state(dims, rng, amp=1.2)
    state = {"rho": st["rho"].astype(np.float64), "v": st["v"][:, :3].astype(np.float64),
             "psi": st["psi"][:, :3].astype(np.float64), "psi_inv": st["psi_inv"][:, :3].astype(np.float64)}
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, alpha=ALPHA)
    a = fr.Adjoint(p, 1, 1)
    a.set_images(pair.target, pair.target_labels, pair.source, pair.source_labels)
    return pair, state, a


def test_step_fwd_is_bit_exact_against_the_fp64_oracle(oracle_mod):
    O = oracle_mod
    dims, k = (22, 19, 17), 3
    n = dims[0] * dims[1] * dims[2]
    say("step fwd", dims, k)
    pair, state, a = _step_case(dims, 21)
    w_reg, tau = float(np.float32(0.2)), float(np.float32(0.2))
    try:
        got = a.step_fwd(k, state, w_reg, tau)
    finally:
        a.close()
    o = O.Oracle(dims, "f64")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, 1.0)

    def four(v):
        out = np.zeros((n, 4))
        out[:, :3] = v
        return out

    o.set_state(v=four(state["v"]), rho=state["rho"], psi=four(state["psi"]), psi_inv=four(state["psi_inv"]))
    t = ALPHA
    for _ in range(k - 1):
        t += ALPHA
    o.step(t, ALPHA, tau, w_reg, 2.0, 0.25)
    ref = o.get_state()
    for key in ("v", "psi", "psi_inv"):
        assert np.array_equal(got[key], ref[key][:, :3]), key
    assert np.array_equal(got["rho"], ref["rho"])


def test_step_bwd_matches_autograd():
    from oracle import unrolled as U
    dims, k = (20, 18, 16), 2
    X, Y, Z = dims
    n = X * Y * Z
    say("step bwd", dims, k)
    pair, state, a = _step_case(dims, 22)
    w_reg, tau = float(np.float32(0.2)), float(np.float32(0.2))
    rng = np.random.default_rng(5)
    seeds = {"rho": rng.normal(size=n), "v": rng.normal(size=(n, 3)), "psi": rng.normal(size=(n, 3)),
             "psi_inv": rng.normal(size=(n, 3))}
    try:
        got, gw, gt = a.step_bwd(k, state, seeds, w_reg, tau)
    finally:
        a.close()
    g = U._Grid(dims)
    T = torch.tensor(pair.target.reshape(n).astype(np.float64))
    S = torch.tensor(pair.source.reshape(n).astype(np.float64))
    I = g.I
    gS = g.scatter(I, torch.stack([(S[I + 1] - S[I - 1]) / 2.0, (S[I + g.sy] - S[I - g.sy]) / 2.0,
                                   (S[I + g.sz] - S[I - g.sz]) / 2.0], dim=1), (n, 3))
    ts = {key: torch.tensor(val, requires_grad=True) for key, val in state.items()}
    wt = torch.tensor(w_reg, dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(tau, dtype=torch.float64, requires_grad=True)
    m1t, m2 = U.schedule(k, ALPHA, 2.0, 0.25)[k - 1]
    out = U.iterate(g, (ts["rho"], ts["v"], ts["psi"], ts["psi_inv"]), T, S, gS, ALPHA, m1t, m2, tt, wt)
    loss = sum((o_ * torch.tensor(seeds[key])).sum() for o_, key in zip(out, ("rho", "v", "psi", "psi_inv")))
    loss.backward()
    for key in ("rho", "v", "psi", "psi_inv"):
        r = _rel(got[key], ts[key].grad.numpy())
        say(key, r)
        assert r <= 1e-12, key
    assert abs(gw - wt.grad.item()) <= 1e-11 * abs(wt.grad.item())
    assert abs(gt - tt.grad.item()) <= 1e-11 * abs(tt.grad.item())
