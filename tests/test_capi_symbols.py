"""The C-ABI libraries load and export every symbol include/mat_gpu.h
declares; the pure-host helpers keep libmat's semantics.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import libmat_amd

HDR = os.path.join(libmat_amd.ROOT, "include", "mat_gpu.h")
pytestmark = pytest.mark.skipif(not libmat_amd.built(), reason="library not built")


def declared():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(mat_\w+)\s*\(", text, flags=re.M)))


def test_header_lists_the_hot_path():
    names = declared()
    for must in ("mat_mul", "mat_rmul", "mat_gemm", "mat_lu", "mat_plu", "mat_chol", "mat_solve",
                 "mat_solve_spd", "mat_solve_tril", "mat_solve_tril_unit", "mat_solve_triu",
                 "mat_solve_trilt", "mat_gpu_gesv_mixed", "mat_gpu_posv_mixed",
                 "mat_gpu_batch_create_f32"):
        assert must in names
    assert set(names) == set(libmat_amd.DROPIN_SYMBOLS + libmat_amd.EXT_SYMBOLS)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_exports(prec):
    lib = C.CDLL(libmat_amd.lib_path(prec))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.mat_gpu_elem_size.restype = C.c_int
    assert lib.mat_gpu_elem_size() == (4 if prec == "f32" else 8)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_host_helpers(prec):
    g = libmat_amd.load(prec)
    L = g.lib
    M = C.POINTER(g.Mat)
    L.mat_from.restype = M
    L.mat_from.argtypes = [C.c_size_t, C.c_size_t, C.c_void_p]
    L.mat_at.restype = g.ct
    L.mat_at.argtypes = [M, C.c_size_t, C.c_size_t]
    vals = np.array([1, 2, 3, 4, 5, 6], dtype=g.dtype)  # row-major initializer (mat.h:1297-1303)
    m = L.mat_from(2, 3, vals.ctypes.data_as(C.c_void_p))
    assert [L.mat_at(m, i, j) for i in range(2) for j in range(3)] == [1, 2, 3, 4, 5, 6]
    stored = np.ctypeslib.as_array(m.contents.data, shape=(6,))
    assert stored.tolist() == [1, 4, 2, 5, 3, 6]  # column-major storage (mat.h:389)
    L.mat_reye.restype = M
    L.mat_reye.argtypes = [C.c_size_t]
    e = L.mat_reye(3)
    assert np.ctypeslib.as_array(e.contents.data, shape=(9,)).tolist() == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    L.mat_equals_tol.restype = C.c_bool
    L.mat_equals_tol.argtypes = [M, M, g.ct]
    L.mat_equals.restype = C.c_bool
    L.mat_equals.argtypes = [M, M]
    assert L.mat_equals(e, e)
    assert not L.mat_equals(e, m)  # shape mismatch
    L.mat_perm.restype = C.POINTER(libmat_amd.Perm)
    L.mat_perm.argtypes = [C.c_size_t]
    L.mat_perm_identity.argtypes = [C.POINTER(libmat_amd.Perm)]
    p = L.mat_perm(4)
    L.mat_perm_identity(p)
    assert [p.contents.data[i] for i in range(4)] == [0, 1, 2, 3]
    L.mat_free_perm.argtypes = [C.POINTER(libmat_amd.Perm)]
    L.mat_free_perm(p)
    L.mat_free_mat(m)
    L.mat_free_mat(e)


@pytest.mark.parametrize("fn", ["mat_plu", "mat_chol", "mat_det"])
def test_empty_matrix_is_a_contract_violation(fn):
    """libmat's MAT_ASSERT_MAT requires rows > 0 and cols > 0 (mat.h:334-340);
    the drop-in aborts with a message on the same violation, before any GPU
    call (so this runs without a GPU), instead of returning a value."""
    import subprocess
    import sys
    code = f"""
import ctypes as C, sys
sys.path.insert(0, {libmat_amd.ROOT!r})
import libmat_amd
g = libmat_amd.load("f32")
buf = (C.c_float * 1)()
z = g.Mat(0, 0, C.cast(buf, C.POINTER(C.c_float)))
pbuf = (C.c_size_t * 1)()
p = libmat_amd.Perm(C.cast(pbuf, C.POINTER(C.c_size_t)), 0)
if "{fn}" == "mat_plu":
    g.lib.mat_plu(C.byref(z), C.byref(z), C.byref(z), C.byref(p))
elif "{fn}" == "mat_chol":
    g.lib.mat_chol(C.byref(z), C.byref(z))
else:
    g.lib.mat_det(C.byref(z))
print("returned")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "assertion failed" in r.stderr and fn in r.stderr


def test_qr_shape_contract():
    """mat_qr asserts m >= n (mat.h:3914); the drop-in aborts with a message
    on a wide matrix before any GPU call, so this runs without a GPU."""
    import subprocess
    import sys
    code = f"""
import ctypes as C, sys
sys.path.insert(0, {libmat_amd.ROOT!r})
import libmat_amd
g = libmat_amd.load("f32")
buf = (C.c_float * 64)()
a = g.Mat(2, 3, C.cast(buf, C.POINTER(C.c_float)))
q = g.Mat(2, 2, C.cast(buf, C.POINTER(C.c_float)))
g.lib.mat_qr(C.byref(a), C.byref(q), C.byref(a))
print("returned")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "assertion failed" in r.stderr and "mat_qr" in r.stderr


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_geqrf_rejects_bad_shapes(prec):
    """The device entry point checks its shapes on the host before it touches
    the device: m < n, n = 0 and lda < m return a negative code with a
    message (no GPU needed)."""
    g = libmat_amd.load(prec)
    fn = getattr(g.lib, f"mat_gpu_geqrf_{prec}")
    g.lib.mat_gpu_last_error.restype = C.c_char_p
    for m, n, lda, what in ((2, 3, 2, b"m >= n"), (3, 0, 3, b"m >= n >= 1"), (4, 2, 3, b"lda")):
        assert fn(None, m, n, None, lda, None) < 0, (m, n, lda)
        assert what in g.lib.mat_gpu_last_error(), (m, n, lda)


def test_batch_rejects_bad_orders():
    """The batched API checks every matrix order (1..64) on the host before
    it allocates or touches the device: a plan with an order outside the
    range returns a negative code with a message and hands out no plan
    (no GPU needed)."""
    g = libmat_amd.load("f32")
    g.lib.mat_gpu_last_error.restype = C.c_char_p
    for bad in (0, 65):
        n = np.array([3, bad, 8], np.int32)
        kind = np.zeros(3, np.int8)
        plan = C.c_void_p(0)
        rc = g.lib.mat_gpu_batch_create_f32(None, C.c_size_t(3), n.ctypes.data_as(C.c_void_p),
                                            kind.ctypes.data_as(C.c_void_p), None, None, None,
                                            None, None, None, 0, C.byref(plan))
        assert rc < 0 and plan.value is None, bad
        assert b"matrix order" in g.lib.mat_gpu_last_error()
        rc = g.lib.mat_gpu_batch_solve_f32(None, C.c_size_t(3), n.ctypes.data_as(C.c_void_p),
                                           kind.ctypes.data_as(C.c_void_p), None, None, None,
                                           None, None, None)
        assert rc < 0 and b"matrix order" in g.lib.mat_gpu_last_error()
