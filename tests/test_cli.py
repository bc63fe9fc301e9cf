"""The fastreg_app CLI (the reference's src/apps/demo.cpp): argument handling, params.ini and
NIfTI I/O on the CPU; a full registration through the app on the GPU, checked against the
library call and the oracle."""
import os
import subprocess

import numpy as np
import pytest

import fastreg_amd as fr
from fastreg_amd import nifti, phantoms

APP = os.path.join(fr.PKG_DIR, "fastreg_app")

INI = """VOL_DIMS_X={X}
VOL_DIMS_Y={Y}
VOL_DIMS_Z={Z}
VOL_SIZE_X=1.0
VOL_SIZE_Y=1.0
VOL_SIZE_Z=1.0
NO_LABELS={L}
ALPHA=0.2
MAX_ITER=15
MAX_UPDATE_NORM=-1e-10
RHO_0=1.0
TAU=0.2
W_REG=0.2
RESET_ITER=-1
p=2.0
C=0.25
"""


def _inputs(tmp_path, shape=(20, 24, 28), seed=5):
    pair = phantoms.blobs(shape=shape, seed=seed, n_blobs=6, max_v=1.5, steps=4)
    X, Y, Z = pair.dims
    paths = {}
    for name, arr in (("tgt", pair.target), ("tgt_lab", pair.target_labels), ("src", pair.source),
                      ("src_lab", pair.source_labels)):
        paths[name] = str(tmp_path / f"{name}.nii")
        nifti.write(paths[name], arr)
    paths["ini"] = str(tmp_path / "params.ini")
    open(paths["ini"], "w").write(INI.format(X=X, Y=Y, Z=Z, L=pair.no_labels))
    out = tmp_path / "out"
    out.mkdir()
    paths["out"] = str(out)
    return pair, paths


def test_nifti_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    for dt in (np.float32, np.int16, np.uint8, np.float64):
        a = (rng.random((5, 6, 7)) * 100).astype(dt)
        p = str(tmp_path / f"a_{np.dtype(dt).name}.nii")
        nifti.write(p, a)
        b = nifti.read(p)
        assert b.dtype == a.dtype and np.array_equal(a, b)


def test_app_help_and_argument_errors():
    r = subprocess.run([APP, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("USAGE: fastreg")
    r = subprocess.run([APP, "a", "b"], capture_output=True, text=True)
    assert r.returncode != 0 and "error: incorrect number of arguments" in r.stderr


def test_app_rejects_bad_ini_and_missing_files(tmp_path):
    bad = tmp_path / "bad.ini"
    bad.write_text("BOGUS=1\n")
    r = subprocess.run([APP, str(bad), "t", "tl", "s", "sl", str(tmp_path)], capture_output=True, text=True)
    assert r.returncode != 0 and "unrecognised option 'BOGUS'" in r.stderr
    good = tmp_path / "good.ini"
    good.write_text("VOL_DIMS_X=4\nVOL_DIMS_Y=4\nVOL_DIMS_Z=4\n")
    r = subprocess.run([APP, str(good), str(tmp_path / "nope.nii"), "tl", "s", "sl", str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "cannot open" in r.stderr


def test_app_rejects_dimension_mismatch(tmp_path):
    pair, p = _inputs(tmp_path)
    open(p["ini"], "w").write(INI.format(X=10, Y=10, Z=10, L=pair.no_labels))
    r = subprocess.run([APP, p["ini"], p["tgt"], p["tgt_lab"], p["src"], p["src_lab"], p["out"]],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "do not match VOL_DIMS" in r.stderr


@pytest.mark.gpu
def test_app_registration_matches_library_and_oracle(tmp_path, oracle_mod):
    O = oracle_mod
    pair, p = _inputs(tmp_path)
    r = subprocess.run([APP, "--verbose", p["ini"], p["tgt"], p["tgt_lab"], p["src"], p["src_lab"], p["out"]],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "RUNNING NON-RIGID REGISTRATION" in r.stdout and "--- SAVED OUTPUT" in r.stdout
    assert "SOLVER REACHED MAX. NO. OF ITERATIONS WITHOUT CONVERGING" in r.stdout
    assert "DICE COEFFICIENT" in r.stdout and "all labels:" in r.stdout
    name = "src.nii"
    outs = [nifti.read(os.path.join(p["out"], f)) for f in
            (f"target_psi_inv_{name}.nii", f"target_psi_inv_labels_{name}.nii", f"source_psi_{name}.nii",
             f"source_psi_labels_{name}.nii")]
    o = O.Oracle(pair.dims, "f32")
    o.init(pair.target, pair.target_labels, pair.source, pair.source_labels)
    log = o.run(O.SolverParams(max_iter=15, alpha=0.2, no_labels=pair.no_labels, verbosity=1))
    for a, b in zip(outs, o.outputs()):
        assert np.array_equal(a.ravel(), b)


def _header_file(path, dims, vox_offset=352.0, datatype=16):
    """A NIfTI-1 header (348 bytes + 4-byte extension flag) with the given dims, and no data."""
    import struct
    h = bytearray(352)
    struct.pack_into("<i", h, 0, 348)
    struct.pack_into("<8h", h, 40, 3, *dims, 1, 1, 1, 1)
    struct.pack_into("<h", h, 70, datatype)
    struct.pack_into("<f", h, 108, vox_offset)
    h[344:348] = b"n+1\0"
    open(path, "wb").write(bytes(h))


def test_app_rejects_malformed_nifti_headers(tmp_path):
    """Huge dimensions, a non-finite or out-of-range vox_offset and missing voxel data end in an
    error message and a nonzero status, not an allocation failure or a crash."""
    good = tmp_path / "good.ini"
    good.write_text("VOL_DIMS_X=4\nVOL_DIMS_Y=4\nVOL_DIMS_Z=4\n")
    cases = {"huge.nii": dict(dims=(32000, 32000, 32000)), "nan_off.nii": dict(dims=(4, 4, 4), vox_offset=float("nan")),
             "neg_off.nii": dict(dims=(4, 4, 4), vox_offset=-8.0), "short.nii": dict(dims=(4, 4, 4))}
    expect = {"huge.nii": "too large", "nan_off.nii": "bad vox_offset", "neg_off.nii": "bad vox_offset",
              "short.nii": "truncated voxel data"}
    for name, kw in cases.items():
        path = str(tmp_path / name)
        _header_file(path, **kw)
        r = subprocess.run([APP, str(good), path, "tl", "s", "sl", str(tmp_path)], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode not in (0, -6, -11, 134, 139), (name, r.returncode, r.stderr)
        assert expect[name] in r.stderr, (name, r.stderr)
