"""C-ABI checks that need no GPU: the library loads, exports every function include/fastreg_c.h
declares, and its host-side logic (parameters, params.ini parsing, argument validation) behaves
like the reference's (src/apps/demo.cpp:96-149, include/params.hpp:7-26, params.ini:1-25)."""
import ctypes as C
import subprocess

import numpy as np
import pytest

import fastreg_amd as fr


def test_library_exports_every_declared_function():
    L = fr.lib()
    names = fr.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", fr.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_default_params_are_the_reference_defaults():
    p = fr.default_params()
    assert list(p.vol_dims) == [193, 229, 193]
    assert list(p.vol_size) == [1.0, 1.0, 1.0]
    assert p.no_labels == 6 and p.max_iter == 2048 and p.reset_iter == -1 and p.verbosity == 0
    assert p.max_update_norm == np.float32(-1e-10)
    assert (p.alpha, p.rho_0, p.tau, p.w_reg, p.p, p.C) == tuple(np.float32(x) for x in (0.004, 1.0, 0.2, 0.2, 2.0, 0.25))


INI = """# IMAGES
VOL_DIMS_X=182
VOL_DIMS_Y=218
VOL_DIMS_Z=182

VOL_SIZE_X=1.0
VOL_SIZE_Y=1.0
VOL_SIZE_Z=1.0

NO_LABELS=33

# SOLVER
ALPHA=0.004
MAX_ITER=300
MAX_UPDATE_NORM=-1e-10
RHO_0=1.0
TAU=0.2
W_REG=0.2
RESET_ITER=-1
p=2.0
C=0.25
"""


def test_load_ini_reference_keys(tmp_path):
    f = tmp_path / "params.ini"
    f.write_text(INI)
    p = fr.load_ini(f)
    assert list(p.vol_dims) == [182, 218, 182] and p.no_labels == 33 and p.max_iter == 300
    assert p.alpha == np.float32(0.004) and p.C == np.float32(0.25)


def test_load_ini_keeps_unset_keys(tmp_path):
    f = tmp_path / "partial.ini"
    f.write_text("MAX_ITER=7\n  # comment\n\nTAU = 0.5\n")
    p = fr.load_ini(f, fr.default_params(no_labels=11))
    assert p.max_iter == 7 and p.tau == np.float32(0.5) and p.no_labels == 11


@pytest.mark.parametrize("text,code", [("UNKNOWN_KEY=1\n", fr.FR_ERR_PARSE), ("MAX_ITER=abc\n", fr.FR_ERR_PARSE),
                                       ("MAX_ITER\n", fr.FR_ERR_PARSE), ("ALPHA=0.1x\n", fr.FR_ERR_PARSE),
                                       ("alpha=0.1\n", fr.FR_ERR_PARSE)])
def test_load_ini_rejects_bad_input(tmp_path, text, code):
    f = tmp_path / "bad.ini"
    f.write_text(text)
    with pytest.raises(fr.FastRegError) as e:
        fr.load_ini(f)
    assert e.value.code == code


def test_load_ini_missing_file():
    with pytest.raises(fr.FastRegError) as e:
        fr.load_ini("/nonexistent/params.ini")
    assert e.value.code == fr.FR_ERR_IO


@pytest.mark.parametrize("kw", [dict(vol_dims=(2, 10, 10)), dict(no_labels=0), dict(max_iter=-1),
                                dict(verbosity=3), dict(vol_dims=(2000, 2000, 2000)),
                                dict(alpha=float("nan")), dict(tau=float("inf")), dict(w_reg=float("-inf")),
                                dict(rho_0=float("nan")), dict(max_update_norm=float("nan"))])
def test_create_rejects_invalid_params_before_touching_the_gpu(kw):
    p = fr.default_params(**kw)
    h = C.c_void_p()
    rc = fr.lib().fr_create(C.byref(p), C.byref(h))
    assert rc == fr.FR_ERR_INVALID and not h.value
    assert fr.lib().fr_last_error()


def test_null_arguments_are_errors_not_crashes():
    L = fr.lib()
    assert L.fr_create(None, None) == fr.FR_ERR_INVALID
    assert L.fr_register(None, None, None, None, None) == fr.FR_ERR_INVALID
    assert L.fr_solver_step(None, C.c_float(0.1), None, None) == fr.FR_ERR_INVALID
    assert L.fr_destroy(None) == fr.FR_OK


@pytest.mark.parametrize("K,ckpt,kw", [(0, 1, {}), (8, 0, {}), (8, 9, {}), (4, 2, dict(vol_dims=(2, 8, 8))),
                                       (4, 2, dict(no_labels=0)), (4, 2, dict(alpha=float("nan"))),
                                       (4, 2, dict(vol_dims=(700, 700, 700)))])
def test_adjoint_create_rejects_invalid_arguments_before_touching_the_gpu(K, ckpt, kw):
    p = fr.default_params(**kw)
    h = C.c_void_p()
    rc = fr.lib().fr_adjoint_create(C.byref(p), K, ckpt, C.byref(h))
    assert rc == fr.FR_ERR_INVALID and not h.value
    assert fr.lib().fr_last_error()


def test_adjoint_null_handles_are_errors_not_crashes():
    L = fr.lib()
    d = C.c_double()
    assert L.fr_adjoint_create(None, 4, 2, None) == fr.FR_ERR_INVALID
    assert L.fr_adjoint_forward(None, None, 0.2, 0.2, 0.0, C.byref(d)) == fr.FR_ERR_INVALID
    assert L.fr_adjoint_backward(None, None, None, None) == fr.FR_ERR_INVALID
    assert L.fr_adjoint_get_state(None, 0, None, None, None, None) == fr.FR_ERR_INVALID
    assert L.fr_adjoint_info(None, None, None, None) == fr.FR_ERR_INVALID
    assert L.fr_adjoint_set_images(None, None, None, None, None) == fr.FR_ERR_INVALID
    assert L.fr_adjoint_destroy(None) == fr.FR_OK
