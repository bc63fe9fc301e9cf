"""K unrolled solver iterations and the registration objective in torch FP64, differentiated by
torch autograd — TEST INFRASTRUCTURE ONLY.

This is the reference-path gradient the library's hand-written adjoint (fr_adjoint_* in
include/fastreg_c.h) is checked against: the same iteration as the oracle's FP64 solver
(fr_oracle_impl.h, which follows src/fastreg/cuda/solver.cu:295-444), written as whole-array
torch operations so that autograd, not hand-written code, produces the gradient.  Only tests/
and bench.py's checks import it.

One iteration (state rho, v, psi, psi_inv -> rho', v', psi', psi_inv'):
  rho'     = rho + alpha * div(rho, v)                         interior   solver.cu:101-173
  J        = D psi_inv (central differences)                   interior   vector_fields.cu:503-535
  L        = -Laplacian(psi)                                   interior   vector_fields.cu:275-300
  Lo, Tp   = trilerp(L, psi_inv), trilerp(T, psi_inv)          all        vector_fields.cu:231-247, 81-99
  gU       = ((S - Tp) grad S + w_reg Lo) det_ref(J)           interior   solver.cu:56-80
  dvv, dlv = (Dv)v, the upwind velocity Laplacian               interior / its range, vector_fields.cu:307-404
  v'       = v - alpha (m1t v + dvv - tau dlv + q gU), q = m2 / (|rho'| + 1e-5f)   interior   solver.cu:188-211
  psi'     = psi + alpha trilerp(v', psi)                      interior   solver.cu:230-247
  psi_inv' = psi_inv - alpha J v'                              interior   solver.cu:263-279
with t_1 = alpha, t_k = t_{k-1} + alpha (reset to alpha when k is a multiple of RESET_ITER),
m1t = (p + 1) / t, m2 = C p^2 t^(p - 2).  Objective of the final state:
  E = 0.5 sum (T - S o psi)^2 + lw * 0.5 sum_{l>=1} sum (1[LT = l] - trilerp(1[LS = l], psi))^2.

Values agree with the FP64 oracle to rounding (the oracle's two FMAs, in lerp and the density
update, are a multiply and an add here); tests pin both this forward pass against the oracle and
this gradient against central differences of the oracle.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def schedule(K, alpha, p, C, reset_iter=-1):
    """[(m1t, m2)] for iterations 1..K, in double (solver.cu:198-207, 423-428)."""
    out = []
    t = float(alpha)
    for k in range(1, K + 1):
        if k >= 2:
            t += alpha
            if reset_iter > 0 and k % reset_iter == 0:
                t = float(alpha)
        out.append(((p + 1.0) / t, C * math.pow(p, 2.0) * math.pow(t, p - 2.0)))
    return out


class _Grid:
    def __init__(self, dims, device="cpu"):
        X, Y, Z = (int(d) for d in dims)
        self.X, self.Y, self.Z, self.n = X, Y, Z, X * Y * Z
        idx = np.arange(self.n)
        x, y, z = idx % X, (idx // X) % Y, idx // (X * Y)
        inner = (x >= 1) & (x <= X - 2) & (y >= 1) & (y <= Y - 2) & (z >= 1) & (z <= Z - 2)
        self.device = device
        self.I = torch.from_numpy(idx[inner]).to(device)
        self.sx, self.sy, self.sz = 1, X, X * Y
        lap = (x >= 2) & (x <= X - 3) & (y >= 2) & (y <= Y - 2) & (z >= 2) & (z <= Z - 3)
        self.IL = torch.from_numpy(idx[lap]).to(device)

    def scatter(self, idx, vals, like_shape):
        out = torch.zeros(like_shape, dtype=torch.float64, device=self.device)
        return out.index_put((idx,), vals)


def _lerp(v0, v1, t):
    return t * v0 + (1 - t) * v1


def trilerp(g: _Grid, f, pos):
    """Trilinear sample of f ((N,) or (N, C)) at pos (M, 3), positions clamped into the grid
    (cuda_utils.hpp:60-136)."""
    X, Y, Z = g.X, g.Y, g.Z
    cx = pos[:, 0].clamp(0, X - 1)
    cy = pos[:, 1].clamp(0, Y - 1)
    cz = pos[:, 2].clamp(0, Z - 1)
    gx, gy, gz = (torch.floor(c.detach()).long() for c in (cx, cy, cz))
    a, b, c = cx - gx, cy - gy, cz - gz
    hx = torch.clamp(gx + 1, max=X - 1)
    hy = torch.clamp(gy + 1, max=Y - 1)
    hz = torch.clamp(gz + 1, max=Z - 1)

    def at(xx, yy, zz):
        return f[xx + X * (yy + Y * zz)]

    if f.dim() == 2:
        a, b, c = a[:, None], b[:, None], c[:, None]
    c11 = _lerp(at(hx, hy, hz), at(hx, hy, gz), c)
    c10 = _lerp(at(hx, gy, hz), at(hx, gy, gz), c)
    c01 = _lerp(at(gx, hy, hz), at(gx, hy, gz), c)
    c00 = _lerp(at(gx, gy, hz), at(gx, gy, gz), c)
    return _lerp(_lerp(c11, c10, b), _lerp(c01, c00, b), a)


def _det_ref(m):
    """The reference's det() (cuda_utils.hpp:389-396), not the determinant; m (M, 3, 3)."""
    a = (m[:, 0, 0] * m[:, 1, 1]) * m[:, 2, 2] + (m[:, 1, 0] * m[:, 2, 1]) * m[:, 0, 2] + (m[:, 2, 0] * m[:, 0, 1]) * m[:, 1, 2]
    b = (m[:, 0, 2] * m[:, 1, 1] + m[:, 2, 0]) * ((m[:, 1, 2] * m[:, 2, 1]) * m[:, 0, 0]) + (m[:, 2, 2] * m[:, 0, 1]) * m[:, 1, 0]
    return a - b


def iterate(g: _Grid, st, T, S, gS, alpha, m1t, m2, tau, w_reg):
    """One iteration; st = (rho (N,), v (N,3), psi (N,3), psi_inv (N,3))."""
    rho, v, psi, pinv = st
    I, sx, sy, sz = g.I, g.sx, g.sy, g.sz
    n = g.n
    # density (upwind fluxes, summed left to right)
    p = rho[I]
    vx, vy, vz = v[I, 0], v[I, 1], v[I, 2]
    t11 = torch.where(vx > 0, -vx * p, -vx * rho[I + sx])
    vxm = v[I - sx, 0]
    t12 = torch.where(vxm > 0, vxm * rho[I - sx], vxm * p)
    t21 = torch.where(vy > 0, -vy * p, -vy * rho[I + sy])
    vym = v[I - sy, 1]
    t22 = torch.where(vym > 0, vym * rho[I - sy], vym * p)
    t31 = torch.where(vz > 0, -vz * p, -vz * rho[I + sz])
    vzm = v[I - sz, 2]
    t32 = torch.where(vzm > 0, vzm * rho[I - sz], vzm * p)
    div = ((((t11 + t12) + t21) + t22) + t31) + t32
    rho1 = rho.index_put((I,), alpha * div + p)
    # J = D psi_inv: rows = components, columns = d/dx, d/dy, d/dz
    J = torch.stack([(pinv[I + sx] - pinv[I - sx]) * 0.5, (pinv[I + sy] - pinv[I - sy]) * 0.5,
                     (pinv[I + sz] - pinv[I - sz]) * 0.5], dim=2)  # (M, comp, axis)
    # L = -Laplacian(psi), 0 off the interior
    s = psi[I] * -6.0
    for o in (sx, -sx, sy, -sy, sz, -sz):
        s = s + psi[I + o]
    L = g.scatter(I, s * -1.0, (n, 3))
    Lo = trilerp(g, L, pinv)
    Tp = trilerp(g, T, pinv)
    # potential gradient
    det = _det_ref(J)
    diff = S[I] - Tp[I]
    gU_I = (gS[I] * diff[:, None] + Lo[I] * w_reg) * det[:, None]
    # (Dv)v
    c = v[I]
    xp, xm, yp, ym, zp, zm = v[I + sx], v[I - sx], v[I + sy], v[I - sy], v[I + sz], v[I - sz]
    mx = lambda q: torch.clamp(q, min=0)  # noqa: E731
    mn = lambda q: torch.clamp(q, max=0)  # noqa: E731

    def diag(cc, nb):
        return ((mx(cc) * mx(cc) - mn(cc) * mn(cc)) + mn(nb) * mn(nb)) - mx(nb) * mx(nb)

    def up(w, cc, m_, p_):
        return torch.where(w > 0, w * (cc - m_), w * (p_ - cc))

    ax1 = diag(c[:, 0], xp[:, 0])
    ax2 = up(c[:, 1], c[:, 0], ym[:, 0], yp[:, 0])
    ax3 = up(c[:, 2], c[:, 0], zm[:, 0], zp[:, 0])
    ay1 = up(c[:, 0], c[:, 1], xm[:, 1], xp[:, 1])
    ay2 = diag(c[:, 1], yp[:, 1])
    ay3 = up(c[:, 2], c[:, 1], zm[:, 1], zp[:, 1])
    az1 = up(c[:, 0], c[:, 2], xm[:, 2], xp[:, 2])
    az2 = up(c[:, 1], c[:, 2], ym[:, 2], yp[:, 2])
    az3 = diag(c[:, 2], zp[:, 2])
    dvv_I = torch.stack([(0.5 * ax1 + ax2) + ax3, (ay1 + 0.5 * ay2) + ay3, (az1 + az2) + 0.5 * az3], dim=1)
    # velocity Laplacian on its range, with the reference's y-1 / z-1 indices
    IL = g.IL
    cl = v[IL]

    def second(comp, s1, s_minus_only):
        neg = (cl[:, comp] < 0)[:, None]
        far = torch.where(neg, v[IL + 2 * s1], v[IL - 2 * s1])
        near = v[IL - s1] if s_minus_only else torch.where(neg, v[IL + s1], v[IL - s1])
        return torch.where(neg, (far - 2.0 * near) + cl, (cl - 2.0 * near) + far)

    dl = -1.0 * ((second(0, sx, False) + second(1, sy, True)) + second(2, sz, True))
    dlv = g.scatter(IL, dl, (n, 3))
    # velocity update
    q = m2 / (torch.abs(rho1[I]) + float(np.float32(1e-5)))
    upd = ((((c * m1t) + dvv_I) - dlv[I] * tau) + gU_I * q[:, None]) * alpha
    v1 = v.index_put((I,), c - upd)
    # psi and psi_inv updates
    u = trilerp(g, v1, psi[I]) * alpha
    psi1 = psi.index_put((I,), psi[I] + u)
    Jv = torch.stack([(J[:, 0, 0] * v1[I, 0] + J[:, 0, 1] * v1[I, 1]) + J[:, 0, 2] * v1[I, 2],
                      (J[:, 1, 0] * v1[I, 0] + J[:, 1, 1] * v1[I, 1]) + J[:, 1, 2] * v1[I, 2],
                      (J[:, 2, 0] * v1[I, 0] + J[:, 2, 1] * v1[I, 1]) + J[:, 2, 2] * v1[I, 2]], dim=1)
    pinv1 = pinv.index_put((I,), pinv[I] - Jv * alpha)
    return rho1, v1, psi1, pinv1


def objective(g: _Grid, psi, T, S, LT, LS, no_labels, lw):
    r = T - trilerp(g, S, psi)
    e = 0.5 * (r * r).sum()
    if lw != 0:
        for lab in range(1, no_labels):
            ind_s = (LS == lab).to(torch.float64)
            rl = (LT == lab).to(torch.float64) - trilerp(g, ind_s, psi)
            e = e + lw * 0.5 * (rl * rl).sum()
    return e


def unrolled(dims, tgt, tgt_lab, src, src_lab, v0, w_reg, tau, K, alpha=0.004, rho_0=1.0, p=2.0, C=0.25,
             reset_iter=-1, no_labels=1, label_weight=0.0, keep_states=False, checkpoint=False):
    """E after K iterations from (rho_0, v0, identity, identity).  tgt/src: float32 arrays of N
    voxels (x-fastest); labels int16; v0 (N, 3) float64 tensor (its device is used throughout);
    w_reg, tau float64 tensors or floats.  checkpoint: recompute each iteration in the backward
    pass (torch.utils.checkpoint) instead of keeping its graph.  Returns E (and the list of states
    after each iteration if keep_states)."""
    dev = v0.device
    g = _Grid(dims, dev)
    n = g.n

    def load(a, dt, npdt):
        return torch.from_numpy(np.asarray(a, dt).reshape(n).astype(npdt)).to(dev)

    T = load(tgt, np.float32, np.float64)
    S = load(src, np.float32, np.float64)
    LT = load(tgt_lab, np.int16, np.int64)
    LS = load(src_lab, np.int16, np.int64)
    I = g.I
    gS = g.scatter(I, torch.stack([(S[I + 1] - S[I - 1]) / 2.0, (S[I + g.sy] - S[I - g.sy]) / 2.0,
                                   (S[I + g.sz] - S[I - g.sz]) / 2.0], dim=1), (n, 3))
    idx = np.arange(n)
    ident = torch.from_numpy(np.stack([idx % g.X, (idx // g.X) % g.Y, idx // (g.X * g.Y)], 1).astype(np.float64))
    ident = ident.to(dev)
    st = (torch.full((n,), float(rho_0), dtype=torch.float64, device=dev), v0, ident.clone(), ident.clone())
    states = []
    for m1t, m2 in schedule(K, float(alpha), float(p), float(C), reset_iter):
        if checkpoint:
            from torch.utils.checkpoint import checkpoint as ckpt

            st = ckpt(lambda r, v, ps, pi, w, t, m1t=m1t, m2=m2: iterate(g, (r, v, ps, pi), T, S, gS, float(alpha),
                                                                         m1t, m2, t, w),
                      *st, w_reg, tau, use_reentrant=False)
        else:
            st = iterate(g, st, T, S, gS, float(alpha), m1t, m2, tau, w_reg)
        if keep_states:
            states.append(st)
    e = objective(g, st[2], T, S, LT, LS, no_labels, label_weight)
    return (e, states) if keep_states else e
