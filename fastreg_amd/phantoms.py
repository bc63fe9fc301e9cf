"""Seeded synthetic registration pairs (numpy/scipy).

The workloads the solver is measured and tested on:

* ``blobs``: a target made of isotropic Gaussian blobs clipped to [0, 1], int16 labels from the
  nearest of a few seeded points (0 where the intensity is below 0.05), and a source made by
  warping the target with a smooth diffeomorphism.  The diffeomorphism is the exponential of a
  stationary velocity built from Gaussian RBFs, integrated by scaling and squaring.  The image
  is warped trilinearly, the labels by nearest neighbour, and Gaussian noise is added to the
  source image.  Defaults: 64^3, 12 blobs, 8 labels, 6 RBFs (sigma 8, max |v| 3 voxels),
  6 squaring steps, noise std 0.01.
* ``brain``: a brain-MRI-sized phantom (182 x 218 x 182 by default) made of nested ellipsoidal
  shells at intensities 0.2 / 0.5 / 0.8 times a smooth bias field in [0.9, 1.1], plus noise of
  std 0.02.  It has 32 int16 labels from the Voronoi cells of seeded points inside the outer
  ellipsoid.  The source is made with 20 RBF centres (sigma 12, max displacement 6 voxels) and
  7 squaring steps.

* ``ramp``: a pair on which the differentiated solver has no branch switch near the evaluation
  point: source = slope * (x + y + z), target = source + shift, labels from the octant halves in
  x and y.  The image force then has one sign everywhere, so from a small positive initial
  velocity every component of v stays positive, psi moves forward and psi_inv backward of the
  grid nodes, and no upwind branch, trilinear cell or clamp changes under a small perturbation.

Arrays are returned in the solver's layout: float32 / int16 with shape (Z, Y, X), so that
``a.ravel()`` is x-fastest.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import ndimage


@dataclass
class Pair:
    target: np.ndarray          # float32 (Z, Y, X)
    target_labels: np.ndarray   # int16 (Z, Y, X)
    source: np.ndarray          # float32 (Z, Y, X)
    source_labels: np.ndarray   # int16 (Z, Y, X)
    no_labels: int              # max label + 1

    @property
    def dims(self):
        """(X, Y, Z)"""
        z, y, x = self.target.shape
        return (x, y, z)


def _grid(shape):
    z, y, x = shape
    return (np.arange(z, dtype=np.float64), np.arange(y, dtype=np.float64), np.arange(x, dtype=np.float64))


def rbf_velocity(shape, rng, n_centres, sigma, max_norm):
    """Stationary velocity (3, Z, Y, X) in voxels, ordered (vx, vy, vz): a sum of Gaussian RBFs
    with random centres and N(0,1) vector weights, scaled so that max |v| = max_norm."""
    zs, ys, xs = _grid(shape)
    v = np.zeros((3,) + tuple(shape), np.float64)
    for _ in range(n_centres):
        c = rng.uniform([0, 0, 0], [shape[2] - 1, shape[1] - 1, shape[0] - 1])
        w = rng.normal(size=3)
        gx = np.exp(-(xs - c[0]) ** 2 / (2 * sigma ** 2))
        gy = np.exp(-(ys - c[1]) ** 2 / (2 * sigma ** 2))
        gz = np.exp(-(zs - c[2]) ** 2 / (2 * sigma ** 2))
        g = gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        for k in range(3):
            v[k] += w[k] * g
    m = np.sqrt((v ** 2).sum(0)).max()
    if m > 0:
        v *= max_norm / m
    return v


def exp_velocity(v, steps):
    """Scaling and squaring: phi = (id + v / 2^steps) composed with itself `steps` times.
    Returns the displacement u (3, Z, Y, X), phi(p) = p + u(p)."""
    shape = v.shape[1:]
    zz, yy, xx = np.meshgrid(*_grid(shape), indexing="ij")
    u = v / (2.0 ** steps)
    for _ in range(steps):
        coords = np.stack([zz + u[2], yy + u[1], xx + u[0]])
        u = u + np.stack([ndimage.map_coordinates(u[k], coords, order=1, mode="nearest") for k in range(3)])
    return u


def warp(img, u, order):
    """img o phi with phi = id + u; order 1 = trilinear intensity, 0 = nearest label."""
    shape = img.shape
    zz, yy, xx = np.meshgrid(*_grid(shape), indexing="ij")
    coords = np.stack([zz + u[2], yy + u[1], xx + u[0]])
    return ndimage.map_coordinates(img, coords, order=order, mode="nearest")


def blobs(shape=(64, 64, 64), seed=1, n_blobs=12, n_labels=8, n_rbf=6, rbf_sigma=8.0, max_v=3.0, steps=6,
          noise=0.01, scale=1.0) -> Pair:
    """The blob phantom pair; `shape` is (Z, Y, X).  `scale` multiplies every length (blob sigmas,
    RBF sigma, max |v|), e.g. 2 for the 64^3 phantom drawn at 128^3."""
    rng = np.random.default_rng(seed)
    zs, ys, xs = _grid(shape)
    t = np.zeros(shape, np.float64)
    for _ in range(n_blobs):
        c = rng.uniform([0, 0, 0], [shape[2] - 1, shape[1] - 1, shape[0] - 1])
        s = rng.uniform(4.0, 10.0) * scale
        a = rng.uniform(0.3, 1.0)
        gx = np.exp(-(xs - c[0]) ** 2 / (2 * s * s))
        gy = np.exp(-(ys - c[1]) ** 2 / (2 * s * s))
        gz = np.exp(-(zs - c[2]) ** 2 / (2 * s * s))
        t += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    t = np.clip(t, 0.0, 1.0)
    pts = rng.uniform([0, 0, 0], [shape[2] - 1, shape[1] - 1, shape[0] - 1], size=(n_labels, 3))
    lab = _voronoi(shape, pts) + 1
    lab[t < 0.05] = 0
    v = rbf_velocity(shape, rng, n_rbf, rbf_sigma * scale, max_v * scale)
    u = exp_velocity(v, steps)
    src = warp(t, u, 1) + rng.normal(0.0, noise, size=shape)
    src_lab = warp(lab.astype(np.float64), u, 0)
    return Pair(t.astype(np.float32), lab.astype(np.int16), src.astype(np.float32),
                np.rint(src_lab).astype(np.int16), n_labels + 1)


def _voronoi(shape, pts):
    """index of the nearest point (x, y, z order) for every voxel, (Z, Y, X) int array"""
    zs, ys, xs = _grid(shape)
    best = np.full(shape, np.inf)
    idx = np.zeros(shape, np.int32)
    for k, p in enumerate(pts):
        d = ((xs - p[0]) ** 2)[None, None, :] + ((ys - p[1]) ** 2)[None, :, None] + ((zs - p[2]) ** 2)[:, None, None]
        m = d < best
        best[m] = d[m]
        idx[m] = k
    return idx


def brain(shape=(182, 218, 182), seed=2, n_labels=32, n_rbf=20, rbf_sigma=12.0, max_v=6.0, steps=7,
          noise=0.02) -> Pair:
    """The brain-sized ellipsoid phantom pair; `shape` is (Z, Y, X)."""
    rng = np.random.default_rng(seed)
    zs, ys, xs = _grid(shape)
    cz, cy, cx = (shape[0] - 1) / 2, (shape[1] - 1) / 2, (shape[2] - 1) / 2
    rz, ry, rx = 0.42 * shape[0], 0.42 * shape[1], 0.42 * shape[2]
    r2 = (((zs - cz) / rz) ** 2)[:, None, None] + (((ys - cy) / ry) ** 2)[None, :, None] + (((xs - cx) / rx) ** 2)[None, None, :]
    t = np.zeros(shape, np.float64)
    t[r2 <= 1.0] = 0.2
    t[r2 <= 0.7 ** 2] = 0.5
    t[r2 <= 0.4 ** 2] = 0.8
    # smooth bias field in [0.9, 1.1]
    ph = rng.uniform(0, 2 * np.pi, 3)
    b = (np.cos(2 * np.pi * zs / shape[0] + ph[0])[:, None, None] + np.cos(2 * np.pi * ys / shape[1] + ph[1])[None, :, None]
         + np.cos(2 * np.pi * xs / shape[2] + ph[2])[None, None, :]) / 3.0
    t = t * (1.0 + 0.1 * b) + rng.normal(0.0, noise, size=shape)
    # label seeds inside the outer ellipsoid
    pts = []
    while len(pts) < n_labels:
        p = rng.uniform([cx - rx, cy - ry, cz - rz], [cx + rx, cy + ry, cz + rz])
        if ((p[0] - cx) / rx) ** 2 + ((p[1] - cy) / ry) ** 2 + ((p[2] - cz) / rz) ** 2 <= 1.0:
            pts.append(p)
    lab = _voronoi(shape, np.array(pts)) + 1
    lab[r2 > 1.0] = 0
    v = rbf_velocity(shape, rng, n_rbf, rbf_sigma, max_v)
    u = exp_velocity(v, steps)
    src = warp(t, u, 1)
    src_lab = warp(lab.astype(np.float64), u, 0)
    return Pair(t.astype(np.float32), lab.astype(np.int16), src.astype(np.float32),
                np.rint(src_lab).astype(np.int16), n_labels + 1)


def random_state(dims, rng, amp=1.5, smooth=2.0):
    """A smooth, in-range solver state for teacher-forced tests: returns dict of float32 arrays
    v (N,4), rho (N,), psi (N,4), psi_inv (N,4) with boundary voxels at their fixed values
    (v = 0, rho = 1, psi = psi_inv = identity)."""
    X, Y, Z = dims
    shape = (Z, Y, X)

    def field(scale):
        f = np.stack([ndimage.gaussian_filter(rng.normal(size=shape), smooth) for _ in range(3)])
        m = np.abs(f).max()
        return f * (scale / (m if m > 0 else 1.0))

    interior = np.zeros(shape, bool)
    interior[1:-1, 1:-1, 1:-1] = True
    zz, yy, xx = np.meshgrid(*_grid(shape), indexing="ij")
    ident = np.stack([xx, yy, zz, np.zeros_like(xx)], -1).reshape(-1, 4)

    def vec(f):
        out = np.zeros((Z, Y, X, 4))
        out[..., 0], out[..., 1], out[..., 2] = f[0], f[1], f[2]
        out[~interior] = 0
        return out.reshape(-1, 4)

    v = vec(field(amp))
    psi = ident + vec(field(amp))
    psi_inv = ident + vec(field(amp))
    rho = 1.0 + 0.2 * ndimage.gaussian_filter(rng.normal(size=shape), smooth)
    rho[~interior] = 1.0
    return {"v": v.astype(np.float32), "rho": rho.reshape(-1).astype(np.float32),
            "psi": psi.astype(np.float32), "psi_inv": psi_inv.astype(np.float32)}


def smooth_velocity(dims, seed, amp=1.0, offset=0.1, sigma=1.5):
    """A smooth random velocity (N, 3) float64 for the differentiated solver, x-fastest: each
    component Gaussian-filtered N(0, 1) noise scaled to max |.| = amp, plus `offset` (so that no
    component sits at 0, where the upwind schemes switch branch)."""
    X, Y, Z = dims
    rng = np.random.default_rng(seed)
    f = np.stack([ndimage.gaussian_filter(rng.normal(size=(Z, Y, X)), sigma) for _ in range(3)], -1).reshape(-1, 3)
    return f * (amp / np.abs(f).max()) + offset


def ramp(shape=(128, 128, 128), slope=1.0, shift=1.0):
    """The ramp pair (see the module docstring); 3 labels (1..3)."""
    z, y, x = np.meshgrid(*(np.arange(k) for k in shape), indexing="ij")
    Z, Y, X = shape
    src = (slope * (x + y + z)).astype(np.float32)
    tgt = (src + np.float32(shift)).astype(np.float32)
    lab = (1 + (x >= X // 2) + (y >= Y // 2)).astype(np.int16)
    return Pair(tgt, lab, src, lab.copy(), 4)


def ramp_velocity(dims, seed, base=5e-4, spread=2e-4):
    """A small positive initial velocity for the ramp pair: base + spread * U(0, 1) per component,
    (N, 3) float64, x-fastest."""
    X, Y, Z = dims
    return base + spread * np.random.default_rng(seed).random((X * Y * Z, 3))
