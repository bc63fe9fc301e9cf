"""The reference-path gradient (oracle/unrolled.py: the unrolled iteration in torch FP64,
differentiated by autograd) pinned on the CPU: its forward states against the FP64 oracle, its
warp against the oracle's apply, and its gradient against central finite differences of the
FP64 oracle (a step small enough that no branch of the upwind schemes or of the trilinear cells
changes between the two evaluations)."""
import numpy as np
import pytest
import torch
from scipy import ndimage

from fastreg_amd import phantoms

ALPHA = float(np.float32(0.004))
TAU = float(np.float32(0.2))
W_REG = float(np.float32(0.2))


def _pair(dims, seed):
    X, Y, Z = dims
    return phantoms.blobs(shape=(Z, Y, X), seed=seed, n_blobs=4, n_labels=3, n_rbf=3, rbf_sigma=3.0, max_v=1.0,
                          steps=4)


def _v0(dims, seed, amp=1.0, off=0.1):
    X, Y, Z = dims
    rng = np.random.default_rng(seed)
    f = np.stack([ndimage.gaussian_filter(rng.normal(size=(Z, Y, X)), 1.5) for _ in range(3)], -1)
    f = f.reshape(-1, 3)
    return f * (amp / np.abs(f).max()) + off


def _oracle_states(O, dims, pair, v0, K, w_reg=W_REG, tau=TAU, reset_iter=-1):
    n = dims[0] * dims[1] * dims[2]
    o = O.Oracle(dims, "f64")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels, 1.0)
    v4 = np.zeros((n, 4))
    v4[:, :3] = v0
    o.set_state(v=v4)
    t, out = ALPHA, []
    for k in range(1, K + 1):
        if k >= 2:
            t += ALPHA
            if reset_iter > 0 and k % reset_iter == 0:
                t = ALPHA
        o.step(t, ALPHA, tau, w_reg, 2.0, 0.25)
        out.append(o.get_state())
    return out


@pytest.mark.parametrize("dims,K,reset_iter", [((12, 11, 10), 4, -1), ((9, 13, 7), 5, 2)])
def test_unrolled_forward_matches_the_fp64_oracle(oracle_mod, dims, K, reset_iter):
    from oracle import unrolled as U
    pair = _pair(dims, 3)
    v0 = _v0(dims, 1)
    ref = _oracle_states(oracle_mod, dims, pair, v0, K, reset_iter=reset_iter)
    _, states = U.unrolled(dims, pair.target, pair.target_labels, pair.source, pair.source_labels,
                           torch.tensor(v0), W_REG, TAU, K, alpha=ALPHA, reset_iter=reset_iter,
                           no_labels=pair.no_labels, keep_states=True)
    for k in range(K):
        rho, v, psi, pinv = (a.numpy() for a in states[k])
        o = ref[k]
        # the oracle rounds two FMAs (lerp, density update) once; here they round twice
        np.testing.assert_allclose(rho, o["rho"], rtol=1e-13, atol=0)
        np.testing.assert_allclose(v, o["v"][:, :3], rtol=0, atol=1e-13 * np.abs(o["v"]).max())
        np.testing.assert_allclose(psi, o["psi"][:, :3], rtol=0, atol=1e-12)
        np.testing.assert_allclose(pinv, o["psi_inv"][:, :3], rtol=0, atol=1e-12)


def test_unrolled_warp_matches_the_oracle_apply(oracle_mod):
    from oracle import unrolled as U
    dims = (11, 9, 8)
    n = 11 * 9 * 8
    rng = np.random.default_rng(4)
    img = rng.normal(size=n)
    pos = rng.uniform(-1.5, 12.5, size=(n, 3))
    phi2 = np.zeros((n, 2))
    phi2[:, 0] = img
    psi4 = np.zeros((n, 4))
    psi4[:, :3] = pos
    ref = oracle_mod.apply(dims, phi2, psi4, prec="f64")[:, 0]
    got = U.trilerp(U._Grid(dims), torch.tensor(img), torch.tensor(pos)).numpy()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14)


def test_unrolled_gradient_matches_oracle_finite_differences(oracle_mod):
    from oracle import unrolled as U
    dims = (12, 11, 10)
    n = 12 * 11 * 10
    K, lw = 4, 0.5
    pair = _pair(dims, 3)
    v0 = _v0(dims, 1)
    g_ = U._Grid(dims)
    T = torch.tensor(pair.target.reshape(n).astype(np.float64))
    S = torch.tensor(pair.source.reshape(n).astype(np.float64))
    LT = torch.tensor(pair.target_labels.reshape(n).astype(np.int64))
    LS = torch.tensor(pair.source_labels.reshape(n).astype(np.int64))

    def E_oracle(v, w=W_REG, tau=TAU):
        psi = _oracle_states(oracle_mod, dims, pair, v, K, w, tau)[-1]["psi"][:, :3]
        return U.objective(g_, torch.tensor(psi), T, S, LT, LS, pair.no_labels, lw).item()

    v0t = torch.tensor(v0, requires_grad=True)
    wt = torch.tensor(W_REG, dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    E = U.unrolled(dims, pair.target, pair.target_labels, pair.source, pair.source_labels, v0t, wt, tt, K,
                   alpha=ALPHA, no_labels=pair.no_labels, label_weight=lw)
    assert abs(E.item() - E_oracle(v0)) <= 1e-12 * abs(E.item())
    E.backward()
    g = v0t.grad.numpy()
    eps = 1e-6
    rng = np.random.default_rng(7)
    for _ in range(4):
        d = rng.normal(size=(n, 3))
        fd = (E_oracle(v0 + eps * d) - E_oracle(v0 - eps * d)) / (2 * eps)
        an = float((g * d).sum())
        assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)
    # dE/dW_REG is ~1e-7 of E: a step of 1e-6 would leave FD dominated by the rounding of E
    eps_s = 1e-4
    for name, grad, kw in (("w_reg", wt.grad.item(), "w"), ("tau", tt.grad.item(), "tau")):
        base = dict(w=W_REG, tau=TAU)
        hi, lo = dict(base), dict(base)
        hi[kw] += eps_s
        lo[kw] -= eps_s
        fd = (E_oracle(v0, **hi) - E_oracle(v0, **lo)) / (2 * eps_s)
        assert abs(fd - grad) <= 1e-4 * abs(grad), (name, fd, grad)


def test_unrolled_gradient_matches_finite_differences_at_eps_1e4_on_the_ramp_pair(oracle_mod):
    """On the ramp pair from a small positive v0 no branch switches near the evaluation point
    (every v component stays positive, psi ahead of and psi_inv behind the grid nodes), so central
    differences at the full step of 1e-4 agree with the gradient to ~1e-8 (FD's own truncation).
    The same check on the blob pair fails at this step: it crosses upwind and trilinear kinks."""
    from oracle import unrolled as U
    dims = (12, 11, 10)
    n = 12 * 11 * 10
    K, lw = 16, 1.0
    X, Y, Z = dims
    pair = phantoms.ramp((Z, Y, X))
    v0 = phantoms.ramp_velocity(dims, 2)
    g_ = U._Grid(dims)
    T = torch.tensor(pair.target.reshape(n).astype(np.float64))
    S = torch.tensor(pair.source.reshape(n).astype(np.float64))
    LT = torch.tensor(pair.target_labels.reshape(n).astype(np.int64))
    LS = torch.tensor(pair.source_labels.reshape(n).astype(np.int64))

    def E_oracle(v, w=W_REG, tau=TAU):
        psi = _oracle_states(oracle_mod, dims, pair, v, K, w, tau)[-1]["psi"][:, :3]
        return U.objective(g_, torch.tensor(psi), T, S, LT, LS, pair.no_labels, lw).item()

    v0t = torch.tensor(v0, requires_grad=True)
    wt = torch.tensor(W_REG, dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    E, states = U.unrolled(dims, pair.target, pair.target_labels, pair.source, pair.source_labels, v0t, wt, tt, K,
                           alpha=ALPHA, no_labels=pair.no_labels, label_weight=lw, keep_states=True)
    I = g_.I
    assert min(float(st[1][I].min()) for st in states) > 0  # no velocity component reaches a kink
    E.backward()
    g = v0t.grad.numpy()
    eps = 1e-4

    def close(fd, an):  # 1e-6 relative, plus the rounding of E over the step
        return abs(fd - an) <= 1e-6 * abs(an) + 8 * abs(E.item()) * 2.2e-16 / eps

    rng = np.random.default_rng(9)
    for _ in range(4):
        d = rng.normal(size=(n, 3))
        fd = (E_oracle(v0 + eps * d) - E_oracle(v0 - eps * d)) / (2 * eps)
        an = float((g * d).sum())
        assert close(fd, an), (fd, an)
    for name, grad, kw in (("w_reg", wt.grad.item(), "w"), ("tau", tt.grad.item(), "tau")):
        hi, lo = dict(w=W_REG, tau=TAU), dict(w=W_REG, tau=TAU)
        hi[kw] += eps
        lo[kw] -= eps
        fd = (E_oracle(v0, **hi) - E_oracle(v0, **lo)) / (2 * eps)
        assert close(fd, grad), (name, fd, grad)
