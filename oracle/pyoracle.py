"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the result to compare against.  Never used by
risq_amd/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_P = C.c_void_p
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sig = {
            "or_parity": (C.c_int, [C.c_uint64, C.c_uint64]),
            "or_permute": (C.c_int, [_P, _P]),
            "or_permute_2": (C.c_int, [_P, _P, _P]),
            "or_combine_2": (C.c_uint64, [C.c_int32, C.c_int32]),
            "or_combine_4": (C.c_uint64, [C.c_int32] * 4),
            "or_ham_create": (_P, [C.c_int32, C.c_double, _P, C.c_size_t, _P, C.c_size_t]),
            "or_ham_free": (None, [_P]),
            "or_ham_diag": (None, [_P, _P, C.c_int64, _P]),
            "or_off_diag": (None, [_P, _P, _P, C.c_int64, _P, _P]),
            "or_sparse_create": (_P, [C.c_int64, _P]),
            "or_build_fast_lookup": (C.c_int, [_P, _P, _P, C.c_int64, _P, C.c_int64]),
            "or_build_small": (C.c_int, [_P, _P, _P, C.c_int64, _P, C.c_int64]),
            "or_sparse_n": (C.c_int64, [_P]),
            "or_sparse_nnz": (C.c_int64, [_P]),
            "or_sparse_export": (None, [_P, _P, _P, _P, _P]),
            "or_sparse_free": (None, [_P]),
            "or_spmv": (None, [_P, _P, _P]),
            "or_spmv_csr": (None, [C.c_int64, _P, _P, _P, _P, _P, _P]),
            "or_davidson": (C.c_int, [_P, _P, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double), _P,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "or_davidson_csr": (C.c_int, [C.c_int64, _P, _P, _P, _P, _P, C.c_double, C.c_double, C.c_int,
                                          C.POINTER(C.c_double), _P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "or_sym_eigen": (None, [C.c_int, _P, _P, _P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


def parity(x: int) -> int:
    return lib().or_parity(x & (2**64 - 1), x >> 64)


def _u128(x: int) -> np.ndarray:
    return np.array([x & (2**64 - 1), x >> 64], dtype=np.uint64)


def permute(a: int, b: int) -> int:
    x, y = _u128(a), _u128(b)  # keep the buffers alive across the call
    return lib().or_permute(_p(x), _p(y))


def permute_2(a: int, b: int, v) -> int:
    x, y, w = _u128(a), _u128(b), np.asarray(v, dtype=np.int32)
    return lib().or_permute_2(_p(x), _p(y), _p(w))


class OracleHam:
    def __init__(self, ints):
        self.one = np.ascontiguousarray(ints.one_body, dtype=np.float64)
        self.two = np.ascontiguousarray(ints.two_body, dtype=np.float64)
        self.h = lib().or_ham_create(ints.n_orb, ints.e_core, _p(self.one), self.one.size, _p(self.two),
                                     self.two.size)

    def __del__(self):
        try:
            lib().or_ham_free(self.h)
        except Exception:
            pass

    def ham_diag(self, dets):
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        out = np.empty(dets.shape[0])
        lib().or_ham_diag(self.h, _p(dets), dets.shape[0], _p(out))
        return out

    def off_diag(self, d1, d2):
        d1 = np.ascontiguousarray(d1, dtype=np.uint64)
        d2 = np.ascontiguousarray(d2, dtype=np.uint64)
        m = d1.shape[0]
        val = np.empty(m)
        deg = np.empty(m, dtype=np.int8)
        lib().or_off_diag(self.h, _p(d1), _p(d2), m, _p(val), _p(deg))
        return val, deg


class OracleSparse:
    """The oracle's SparseMatUpperTri, built by the gen_sparse_ham_fast_lookup restatement."""

    def __init__(self, n: int = 0, diag=None):
        diag = np.zeros(0) if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
        self.s = lib().or_sparse_create(n, _p(diag) if n else None)

    def __del__(self):
        try:
            lib().or_sparse_free(self.s)
        except Exception:
            pass

    def build(self, ham: OracleHam, dets, diag, n_stored: int = 0):
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        rc = lib().or_build_fast_lookup(self.s, ham.h, _p(dets), dets.shape[0], _p(diag), n_stored)
        assert rc == 0
        return self

    def build_small(self, ham: OracleHam, dets, diag, n_stored: int = 0):
        """gen_sparse_ham_fast, opp_algo = same_algo = 1 (the n <= 100 build).
        Raises ValueError when the space is not closed under spin flip."""
        dets = np.ascontiguousarray(dets, dtype=np.uint64)
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        rc = lib().or_build_small(self.s, ham.h, _p(dets), dets.shape[0], _p(diag), n_stored)
        if rc == -2:
            raise ValueError("space not closed under spin flip")
        assert rc == 0
        return self

    @property
    def n(self):
        return lib().or_sparse_n(self.s)

    @property
    def nnz(self):
        return lib().or_sparse_nnz(self.s)

    def export(self):
        n, nnz = self.n, self.nnz
        rp = np.empty(n + 1, dtype=np.int64)
        col = np.empty(max(nnz, 1), dtype=np.int32)
        val = np.empty(max(nnz, 1))
        diag = np.empty(n)
        lib().or_sparse_export(self.s, _p(rp), _p(col), _p(val), _p(diag))
        return rp, col[:nnz], val[:nnz], diag

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        lib().or_spmv(self.s, _p(x), _p(y))
        return y

    def davidson(self, init=None, tol=1e-4, etol=1e-8, max_iters=50):
        e0 = C.c_double()
        it = C.c_int()
        ns = C.c_int()
        coeffs = np.empty(self.n)
        if init is not None:
            init = np.ascontiguousarray(init, dtype=np.float64)
        rc = lib().or_davidson(self.s, _p(init), tol, etol, max_iters, C.byref(e0), _p(coeffs), C.byref(it),
                               C.byref(ns))
        return rc, e0.value, coeffs, it.value, ns.value


def spmv_csr(rp, col, val, diag, x):
    n = rp.size - 1
    y = np.empty(n)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().or_spmv_csr(n, _p(rp), _p(col), _p(val), _p(diag), _p(x), _p(y))
    return y


def davidson_csr(rp, col, val, diag, init=None, tol=1e-4, etol=1e-8, max_iters=50):
    n = rp.size - 1
    e0 = C.c_double()
    it = C.c_int()
    ns = C.c_int()
    coeffs = np.empty(n)
    if init is not None:
        init = np.ascontiguousarray(init, dtype=np.float64)
    rc = lib().or_davidson_csr(n, _p(rp), _p(col), _p(val), _p(diag), _p(init), tol, etol, max_iters,
                               C.byref(e0), _p(coeffs), C.byref(it), C.byref(ns))
    return rc, e0.value, coeffs, it.value, ns.value


def sym_eigen(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    k = a.shape[0]
    vals = np.empty(k)
    vecs = np.empty((k, k))
    lib().or_sym_eigen(k, _p(a), _p(vals), _p(vecs))
    return vals, vecs
