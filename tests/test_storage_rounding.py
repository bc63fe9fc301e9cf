"""The oracle's rounding for the 16-bit storage forms of v and rho, pinned against numpy's
float16 and torch's bfloat16 conversions (both IEEE round to nearest even): every finite fp16
value, every midpoint between neighbouring fp16 values (the ties) and the floats next to them,
the overflow and subnormal edges, and random values over 15 decades."""
import numpy as np
import pytest

from oracle import oracle as O


def _cases():
    rng = np.random.default_rng(0)
    n = 400_000
    rand = (rng.standard_normal(n) * 10.0 ** rng.integers(-9, 6, n)).astype(np.float32)
    edges = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, -65520.0, 6.1035e-05, 5.96e-8, 2.98e-8,
                      2.9802322e-08, 1e-30, 3.4e38, np.inf, -np.inf], np.float32)
    h = np.arange(0, 65536, dtype=np.uint16).view(np.float16).astype(np.float32)
    hs = np.unique(h[np.isfinite(h)])
    mid = ((hs[:-1].astype(np.float64) + hs[1:]) / 2).astype(np.float32)
    return np.concatenate([rand, edges, hs, mid, np.nextafter(mid, np.float32(np.inf)),
                           np.nextafter(mid, np.float32(-np.inf))])


def _same(a, b):
    return (a == b) | (np.isnan(a) & np.isnan(b))


def test_fp16_rounding_matches_numpy(oracle_mod):
    a = _cases()
    with np.errstate(over="ignore"):
        ref = a.astype(np.float16).astype(np.float32)
    assert _same(O.round16(O.STORAGE_F16, a), ref).all()


def test_bf16_rounding_matches_torch(oracle_mod):
    torch = pytest.importorskip("torch")
    a = _cases()
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert _same(O.round16(O.STORAGE_BF16, a), ref).all()


def test_fp32_storage_is_identity(oracle_mod):
    a = _cases()
    assert _same(O.round16(O.STORAGE_F32, a), a).all()


def test_quantised_oracle_fixed_point(oracle_mod):
    """S = T stays a bit-exact fixed point in the 16-bit storage forms too (every update is 0)."""
    from fastreg_amd import phantoms
    pair = phantoms.blobs(shape=(12, 14, 16), seed=3)
    for st in (O.STORAGE_BF16, O.STORAGE_F16):
        o = O.Oracle((16, 14, 12), "f32", storage=st)
        o.init(pair.target, pair.target_labels, pair.target, pair.target_labels)
        s0 = o.get_state()
        o.run(O.SolverParams(max_iter=4, alpha=0.3, no_labels=pair.no_labels))
        s1 = o.get_state()
        for k in ("v", "rho", "psi", "psi_inv"):
            assert np.array_equal(s0[k], s1[k]), (st, k)
