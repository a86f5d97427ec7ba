"""libmat_amd: MI355X (gfx950) implementation of libmat's dense GEMM and
factor-and-solve path behind libmat's own C API.

The product is the pair of shared libraries libmatgpu_f32.so /
libmatgpu_f64.so (C ABI, include/mat_gpu.h).  This module is the Python
host-side mirror of that interface: it loads a flavour with ctypes, lays
out libmat's ``Mat`` / ``Perm`` structs (mat.h:364-385) and exposes the
drop-in entry points under the reference's names (``mat_mul``, ``mat_plu``,
``mat_solve``, ...) on numpy arrays, plus the device-resident extension
entry points on raw device pointers.

There is no CPU fallback: if the HIP library is missing, ``load`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# Drop-in symbols (include/mat_gpu.h part 1) and extension symbols (part 2).
DROPIN_SYMBOLS = [
    "mat_mul", "mat_rmul", "mat_gemm", "mat_syrk", "mat_syrk_t", "mat_qr", "mat_qr_r", "mat_lu",
    "mat_plu",
    "mat_chol", "mat_solve",
    "mat_solve_spd", "mat_solve_tril", "mat_solve_tril_unit", "mat_solve_triu",
    "mat_solve_trilt", "mat_det", "mat_inv",
    "mat_empty", "mat_mat", "mat_from", "mat_init", "mat_free_mat", "mat_zeros", "mat_eye",
    "mat_reye", "mat_vec", "mat_vec_from", "mat_perm", "mat_free_perm", "mat_perm_identity",
    "mat_perm_mat", "mat_deep_copy", "mat_rdeep_copy", "mat_at", "mat_set_at", "mat_equals",
    "mat_equals_tol",
]
EXT_SYMBOLS = [
    "mat_gpu_last_error", "mat_gpu_elem_size", "mat_gpu_stream_check",
    "mat_gpu_timing_enable", "mat_gpu_timing_query",
    "mat_gpu_gemm_f32", "mat_gpu_gemm_f64", "mat_gpu_gemm_lowp", "mat_gpu_getrf_f32", "mat_gpu_getrf_f64",
    "mat_gpu_getrf_tournament_f32",
    "mat_gpu_getrs_f32", "mat_gpu_getrs_f64", "mat_gpu_potrf_f32", "mat_gpu_potrf_f64",
    "mat_gpu_geqrf_f32", "mat_gpu_geqrf_f64",
    "mat_gpu_potrs_f32", "mat_gpu_potrs_f64", "mat_gpu_gesv_f32", "mat_gpu_gesv_f64",
    "mat_gpu_posv_f32", "mat_gpu_posv_f64", "mat_gpu_ir_defaults", "mat_gpu_gesv_mixed",
    "mat_gpu_posv_mixed", "mat_gpu_batch_create_f32", "mat_gpu_batch_run",
    "mat_gpu_batch_tuning", "mat_gpu_batch_tuning_shape", "mat_gpu_batch_destroy",
    "mat_gpu_batch_solve_f32",
    "mat_gpu_batch_solve_noalloc_f32",
]


def lib_path(prec: str) -> str:
    return os.path.join(HERE, f"libmatgpu_{prec}.so")


class IrOpts(C.Structure):
    _fields_ = [("passes", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double),
                ("leaf", C.c_int), ("pivoting", C.c_int)]


class IrStatus(C.Structure):
    _fields_ = [("iterations", C.c_int), ("fell_back", C.c_int),
                ("backward_error", C.c_double), ("factor_ms", C.c_double),
                ("total_ms", C.c_double), ("precision", C.c_int)]


class Perm(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_size_t)), ("size", C.c_size_t)]


def _mat_struct(ct):
    class Mat(C.Structure):
        _fields_ = [("rows", C.c_size_t), ("cols", C.c_size_t), ("data", C.POINTER(ct))]
    return Mat


class GpuLib:
    """One flavour of the library (``"f32"`` or ``"f64"``)."""

    def __init__(self, prec: str = "f32"):
        assert prec in ("f32", "f64")
        path = lib_path(prec)
        if not os.path.exists(path):
            raise RuntimeError(f"libmat_amd: {path} is not built "
                               "(run `python -m libmat_amd.build`)")
        self.prec = prec
        self.lib = C.CDLL(path)
        self.dtype = np.float32 if prec == "f32" else np.float64
        self.ct = C.c_float if prec == "f32" else C.c_double
        self.Mat = _mat_struct(self.ct)
        self._sig()

    # ---------------------------------------------------------- signatures
    def _sig(self):
        L, M, P = self.lib, C.POINTER(self.Mat), C.POINTER(Perm)
        vp, sz, i32, dbl, f32 = C.c_void_p, C.c_size_t, C.c_int, C.c_double, C.c_float
        def s(name, res, args):
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        s("mat_mul", None, [M, M, M])
        s("mat_rmul", M, [M, M])
        s("mat_gemm", None, [M, self.ct, M, M, self.ct])
        s("mat_syrk", None, [M, M, self.ct, self.ct, C.c_char])
        s("mat_syrk_t", None, [M, M, self.ct, self.ct, C.c_char])
        s("mat_qr", None, [M, M, M])
        s("mat_qr_r", None, [M, M])
        s("mat_lu", i32, [M, M, M, P, P])
        s("mat_plu", i32, [M, M, M, P])
        s("mat_chol", i32, [M, M])
        for t in ("mat_solve", "mat_solve_tril", "mat_solve_tril_unit", "mat_solve_triu",
                  "mat_solve_trilt"):
            s(t, None, [M, M, M])
        s("mat_solve_spd", i32, [M, M, M])
        s("mat_det", self.ct, [M])
        s("mat_inv", None, [M, M])
        s("mat_free_mat", None, [M])
        s("mat_gpu_last_error", C.c_char_p, [])
        s("mat_gpu_elem_size", i32, [])
        s("mat_gpu_stream_check", i32, [vp])
        s("mat_gpu_timing_enable", i32, [i32])
        s("mat_gpu_timing_query", i32, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(dbl),
                                        C.POINTER(dbl)])
        s("mat_gpu_gemm_f32", i32, [vp, i32, i32, sz, sz, sz, f32, vp, sz, vp, sz, f32, vp, sz])
        s("mat_gpu_gemm_f64", i32, [vp, i32, i32, sz, sz, sz, dbl, vp, sz, vp, sz, dbl, vp, sz])
        s("mat_gpu_gemm_lowp", i32, [vp, i32, sz, sz, sz, f32, vp, sz, vp, sz, vp, sz])
        for t in ("f32", "f64"):
            s(f"mat_gpu_getrf_{t}", i32, [vp, sz, vp, sz, vp, i32, C.POINTER(i32)])
            s(f"mat_gpu_getrs_{t}", i32, [vp, sz, sz, vp, sz, vp, vp, sz, vp, sz])
            s(f"mat_gpu_potrf_{t}", i32, [vp, sz, vp, sz])
            s(f"mat_gpu_geqrf_{t}", i32, [vp, sz, sz, vp, sz, vp])
            s(f"mat_gpu_potrs_{t}", i32, [vp, sz, sz, vp, sz, vp, sz, vp, sz])
            s(f"mat_gpu_gesv_{t}", i32, [vp, sz, sz, vp, sz, vp, sz, vp, sz])
            s(f"mat_gpu_posv_{t}", i32, [vp, sz, sz, vp, sz, vp, sz, vp, sz])
        s("mat_gpu_getrf_tournament_f32", i32, [vp, sz, vp, sz, vp, C.POINTER(i32)])
        s("mat_gpu_ir_defaults", None, [C.POINTER(IrOpts), i32])
        for t in ("mat_gpu_gesv_mixed", "mat_gpu_posv_mixed"):
            s(t, i32, [vp, sz, sz, vp, sz, vp, sz, vp, sz, C.POINTER(IrOpts), C.POINTER(IrStatus)])
        s("mat_gpu_batch_create_f32", i32,
          [vp, sz, vp, vp, vp, vp, vp, vp, vp, vp, i32, C.POINTER(vp)])
        s("mat_gpu_batch_run", i32, [vp, vp])
        s("mat_gpu_batch_tuning", i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(dbl)])
        s("mat_gpu_batch_tuning_shape", i32, [vp, C.POINTER(i32), C.POINTER(C.c_int32),
                                              C.POINTER(dbl)])
        s("mat_gpu_batch_destroy", i32, [vp])
        s("mat_gpu_batch_solve_f32", i32, [vp, sz, vp, vp, vp, vp, vp, vp, vp, vp])
        s("mat_gpu_batch_solve_noalloc_f32", i32, [vp, sz, vp, vp, vp, vp, vp, vp, vp, vp])

    def last_error(self) -> str:
        return self.lib.mat_gpu_last_error().decode()

    def check(self, rc: int, what: str) -> int:
        if rc < -1:
            raise RuntimeError(f"{what} failed ({rc}): {self.last_error()}")
        return rc

    # ------------------------------------------------- Mat marshalling ----
    def to_mat(self, a):
        """numpy (rows, cols) -> (Mat, backing column-major buffer)."""
        a = np.asarray(a, dtype=self.dtype)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        buf = np.asfortranarray(a).copy(order="F")
        m = self.Mat(a.shape[0], a.shape[1], buf.ctypes.data_as(C.POINTER(self.ct)))
        return m, buf

    def empty_mat(self, rows, cols):
        return self.to_mat(np.zeros((rows, cols), self.dtype))

    @staticmethod
    def perm(n):
        buf = np.zeros(n, dtype=np.uintp)
        return Perm(buf.ctypes.data_as(C.POINTER(C.c_size_t)), n), buf

    # ---------------------------------------------- drop-in, numpy level --
    def mat_gemm(self, alpha, A, B, beta, Cm):
        a, ab = self.to_mat(A)
        b, bb = self.to_mat(B)
        c, cb = self.to_mat(Cm)
        self.lib.mat_gemm(C.byref(c), alpha, C.byref(a), C.byref(b), beta)
        return cb

    def mat_syrk(self, alpha, A, beta, Cm, uplo="L", trans=False):
        """mat_syrk (trans: mat_syrk_t) on copies of A and Cm; returns C."""
        a, ab = self.to_mat(A)
        c, cb = self.to_mat(Cm)
        fn = self.lib.mat_syrk_t if trans else self.lib.mat_syrk
        fn(C.byref(c), C.byref(a), alpha, beta, uplo.encode())
        return cb

    def mat_qr(self, A):
        """mat_qr on a copy of A; returns (Q, R)."""
        a, ab = self.to_mat(A)
        q, qb = self.empty_mat(a.rows, a.rows)
        r, rb = self.empty_mat(a.rows, a.cols)
        self.lib.mat_qr(C.byref(a), C.byref(q), C.byref(r))
        return qb, rb

    def mat_qr_r(self, A):
        """mat_qr_r on a copy of A; returns R."""
        a, ab = self.to_mat(A)
        r, rb = self.empty_mat(a.rows, a.cols)
        self.lib.mat_qr_r(C.byref(a), C.byref(r))
        return rb

    def mat_mul(self, A, B):
        a, ab = self.to_mat(A)
        b, bb = self.to_mat(B)
        out, ob = self.empty_mat(a.rows, b.cols)
        self.lib.mat_mul(C.byref(out), C.byref(a), C.byref(b))
        return ob

    def mat_rmul(self, A, B):
        a, ab = self.to_mat(A)
        b, bb = self.to_mat(B)
        r = self.lib.mat_rmul(C.byref(a), C.byref(b))
        m = r.contents
        out = np.ctypeslib.as_array(m.data, shape=(m.rows * m.cols,)).reshape(
            (m.rows, m.cols), order="F").copy()
        self.lib.mat_free_mat(r)
        return out

    def mat_plu(self, A):
        a, ab = self.to_mat(A)
        n = a.rows
        l, lb = self.empty_mat(n, n)
        u, ub = self.empty_mat(n, n)
        p, pb = self.perm(n)
        sw = self.lib.mat_plu(C.byref(a), C.byref(l), C.byref(u), C.byref(p))
        return lb, ub, pb.astype(np.int64), sw

    def mat_lu(self, A):
        a, ab = self.to_mat(A)
        n = a.rows
        l, lb = self.empty_mat(n, n)
        u, ub = self.empty_mat(n, n)
        p, pb = self.perm(n)
        q, qb = self.perm(n)
        sw = self.lib.mat_lu(C.byref(a), C.byref(l), C.byref(u), C.byref(p), C.byref(q))
        return lb, ub, pb.astype(np.int64), qb.astype(np.int64), sw

    def mat_chol(self, A):
        a, ab = self.to_mat(A)
        n = a.rows
        l, lb = self.empty_mat(n, n)
        r = self.lib.mat_chol(C.byref(a), C.byref(l))
        return lb, r

    def _vec_solve(self, fn, T, b):
        t, tb = self.to_mat(T)
        bv, bb = self.to_mat(np.asarray(b).reshape(-1, 1))
        x, xb = self.empty_mat(t.rows, 1)
        r = fn(C.byref(x), C.byref(t), C.byref(bv))
        return xb.reshape(-1), r

    def mat_solve(self, A, b):
        return self._vec_solve(self.lib.mat_solve, A, b)[0]

    def mat_solve_spd(self, A, b):
        return self._vec_solve(self.lib.mat_solve_spd, A, b)

    def mat_solve_tril(self, L, b):
        return self._vec_solve(self.lib.mat_solve_tril, L, b)[0]

    def mat_solve_tril_unit(self, L, b):
        return self._vec_solve(self.lib.mat_solve_tril_unit, L, b)[0]

    def mat_solve_triu(self, U, b):
        return self._vec_solve(self.lib.mat_solve_triu, U, b)[0]

    def mat_solve_trilt(self, L, b):
        return self._vec_solve(self.lib.mat_solve_trilt, L, b)[0]

    def mat_det(self, A):
        a, ab = self.to_mat(A)
        return float(self.lib.mat_det(C.byref(a)))

    def mat_inv(self, A):
        a, ab = self.to_mat(A)
        out, ob = self.empty_mat(a.rows, a.rows)
        self.lib.mat_inv(C.byref(out), C.byref(a))
        return ob


_LIBS: dict[str, GpuLib] = {}


def load(prec: str = "f32") -> GpuLib:
    """Load (once) and return the requested flavour; raises if not built."""
    if prec not in _LIBS:
        _LIBS[prec] = GpuLib(prec)
    return _LIBS[prec]


def built() -> bool:
    return all(os.path.exists(lib_path(p)) for p in ("f32", "f64"))
