"""The reference's further public functions on the GPU (DeformationField::get_inverse,
interpolate_gradient, Reductor::voxel_max_energy, ScalarField::sum) against the oracle."""
import math
import sys

import numpy as np
import pytest

import fastreg_amd as fr
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ident(dims):
    X, Y, Z = dims
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    out = np.zeros((X * Y * Z, 4), np.float32)
    out[:, 0], out[:, 1], out[:, 2] = xx.ravel(), yy.ravel(), zz.ravel()
    return out


@pytest.mark.parametrize("dims,seed", [((23, 19, 17), 0), ((64, 9, 12), 1)])
def test_get_inverse_bit_exact(oracle_mod, dims, seed):
    O = oracle_mod
    say("get_inverse", dims)
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed)
    psi = np.zeros((n, 4), np.float32)
    psi[:, :3] = rng.normal(0, 3.0, (n, 3))
    start = _ident(dims)
    start[:, :3] += rng.normal(0, 1.0, (n, 3)).astype(np.float32)
    g = fr.get_inverse(psi, start, dims)
    o = O.get_inverse(dims, psi, start, 48, "f32")
    assert np.array_equal(g[:, :3], o[:, :3])


@pytest.mark.parametrize("dims,seed", [((23, 19, 17), 2), ((7, 33, 12), 3)])
def test_interpolate_gradient_bit_exact(oracle_mod, dims, seed):
    O = oracle_mod
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 4), np.float32)
    f[:, :3] = rng.normal(size=(n, 3))
    psi = _ident(dims)
    psi[:, :3] += rng.normal(0, 3.0, (n, 3)).astype(np.float32)
    psi[::7, :3] = np.round(psi[::7, :3])  # exact integers: the weight-0 neighbour
    psi[::11, 0] = dims[0] - 1             # clamp edge
    g = fr.interpolate_gradient(f, psi, dims)
    o = O.interpolate_gradient(dims, f, psi, "f32")
    assert np.array_equal(g[:, :3], o[:, :3])


def test_voxel_max_energy_matches_oracle(oracle_mod):
    O = oracle_mod
    dims = (30, 26, 22)
    X, Y, Z = dims
    pair = phantoms.blobs(shape=(Z, Y, X), seed=9)
    p = fr.default_params(vol_dims=dims, no_labels=pair.no_labels, max_iter=10, w_reg=0.3)
    g = fr.Solver(p)
    g.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    rng = np.random.default_rng(9)
    st = phantoms.random_state(dims, rng, amp=2.0)
    g.set_state(**st)
    o.set_state(**st)
    ge, gv = g.voxel_max_energy()
    oe, ov = o.voxel_max_energy(0.3)
    assert (ge, gv) == (oe, ov)


def test_field_sum():
    rng = np.random.default_rng(5)
    for n in (1, 1000, 3_000_001):
        a = rng.normal(size=n).astype(np.float32)
        s = fr.field_sum(a)
        ref = math.fsum(a.astype(np.float64))
        assert abs(s - ref) <= 1e-12 * max(1.0, abs(ref)) + 1e-9, (n, s, ref)
