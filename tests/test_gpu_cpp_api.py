"""The reference's C++ interface on the GPU: tests/cpp/reference_caller (built by build(), written
against the reference's FastReg, Params, fastreg::cuda::Solver, Volume and DeformationField) runs a
registration through FastReg::operator() and through Solver::estimate_psi, twice each (the second
continuing from the state the first left, as the reference's objects do), then exercises
DeformationField::apply and clear and Volume::clear.  Every output is compared with the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

from fastreg_amd import build as B
from fastreg_amd import phantoms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_reference_cpp_interface_matches_oracle(oracle_mod, tmp_path):
    O = oracle_mod
    if not os.path.exists(B.CALLER):
        pytest.fail(f"{B.CALLER} is not built (run build())")
    dims = (34, 30, 26)
    X, Y, Z = dims
    n = X * Y * Z
    iters = 20
    pair = phantoms.blobs(shape=(Z, Y, X), seed=17)
    inp, out = tmp_path / "in", tmp_path / "out"
    inp.mkdir()
    out.mkdir()
    pair.target.astype(np.float32).tofile(inp / "tgt.f32")
    pair.target_labels.astype(np.int16).tofile(inp / "tgt_lab.i16")
    pair.source.astype(np.float32).tofile(inp / "src.f32")
    pair.source_labels.astype(np.int16).tofile(inp / "src_lab.i16")
    say("reference_caller", dims, iters)
    r = subprocess.run([B.CALLER, str(X), str(Y), str(Z), str(iters), str(pair.no_labels), str(inp), str(out)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0 and "reference_caller ok" in r.stdout, r.stdout

    def f32(name, width=1):
        return np.fromfile(out / name, np.float32).reshape((n, width) if width > 1 else n)

    def i16(name):
        return np.fromfile(out / name, np.int16)

    o = O.Oracle(dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o.run(O.SolverParams(max_iter=iters, no_labels=pair.no_labels))
    img1, lab1, img2, lab2 = o.outputs()
    st = o.get_state()
    for stem in ("fastreg", "solver"):
        a1 = f32(f"{stem}_img1") if stem == "fastreg" else f32("solver_img")
        l1 = i16(f"{stem}_lab1") if stem == "fastreg" else i16("solver_lab")
        a2 = f32(f"{stem}_img2") if stem == "fastreg" else f32("solver_n_img")
        l2 = i16(f"{stem}_lab2") if stem == "fastreg" else i16("solver_n_lab")
        assert np.array_equal(a1, img1) and np.array_equal(l1, lab1), stem
        assert np.array_equal(a2, img2) and np.array_equal(l2, lab2), stem
        psi = f32(f"{stem}_psi", 4)
        assert np.array_equal(psi.view(np.uint32), st["psi"].view(np.uint32)), stem
    assert np.array_equal(f32("solver_psi_inv", 4).view(np.uint32), st["psi_inv"].view(np.uint32))
    # the second FastReg call and the second estimate_psi continue from the first's v, psi and
    # psi_inv with rho reset (the reference's state carried between calls); after clear() the
    # solver starts afresh
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    o.set_state(v=st["v"], psi=st["psi"], psi_inv=st["psi_inv"])
    o.run(O.SolverParams(max_iter=iters, no_labels=pair.no_labels))
    img1b, lab1b, img2b, lab2b = o.outputs()
    st2 = o.get_state()
    assert not np.array_equal(st2["psi"], st["psi"])
    for a, b, nm in ((f32("fastreg2_img1"), img1b, "img1"), (i16("fastreg2_lab1"), lab1b, "lab1"),
                     (f32("fastreg2_img2"), img2b, "img2"), (i16("fastreg2_lab2"), lab2b, "lab2"),
                     (f32("solver2_img"), img1b, "solver2 img"), (i16("solver2_lab"), lab1b, "solver2 lab"),
                     (f32("solver2_n_img"), img2b, "solver2 n img"), (i16("solver2_n_lab"), lab2b, "solver2 n lab"),
                     (f32("solver3_img"), img1, "solver3 img"), (i16("solver3_lab"), lab1, "solver3 lab"),
                     (f32("solver3_n_img"), img2, "solver3 n img"), (i16("solver3_n_lab"), lab2, "solver3 n lab")):
        assert np.array_equal(a, b), nm
    for name in ("fastreg2_psi", "solver2_psi"):
        assert np.array_equal(f32(name, 4).view(np.uint32), st2["psi"].view(np.uint32)), name
    # DeformationField::apply with the estimated psi is the source warp of the finalisation
    assert np.array_equal(f32("apply_img"), img2) and np.array_equal(i16("apply_lab"), lab2)
    # clear() makes psi the identity, whose warp is an exact copy
    x, y, z = np.arange(n) % X, (np.arange(n) // X) % Y, np.arange(n) // (X * Y)
    assert np.array_equal(f32("identity_psi", 4), np.stack([x, y, z, 0 * x], 1).astype(np.float32))
    assert np.array_equal(f32("identity_img"), pair.source.reshape(n).astype(np.float32))
    assert np.array_equal(i16("identity_lab"), pair.source_labels.reshape(n).astype(np.int16))
    assert not f32("cleared_img").any() and not i16("cleared_lab").any()
