"""ctypes wrapper of the CPU oracle (oracle/bin/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "bin", "liboracle.so")
CLI = os.path.join(HERE, "bin", "vhci_cpu")

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_L = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sig = {
            "orc_error": (ctypes.c_char_p, []),
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_load": (_vp, [ctypes.c_char_p, ctypes.c_int]),
            "orc_free": (None, [_vp]),
            "orc_info_i": (ctypes.c_int64, [_vp, ctypes.c_int]),
            "orc_info_d": (ctypes.c_double, [_vp, ctypes.c_int]),
            "orc_freqs": (ctypes.c_int, [_vp, _dp]),
            "orc_fc_count": (ctypes.c_int, [_vp, ctypes.c_int]),
            "orc_fc_get": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _ip, _ip, _ip, _ip, _ip]),
            "orc_initial_basis": (ctypes.c_int, [_vp, _i32p]),
            "orc_coeffsum": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_anharm_pot": (ctypes.c_double, [_vp, ctypes.c_int, ctypes.c_int, _i32p, _i32p]),
            "orc_build_h": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _i64p, _i64p, _i32p, _dp]),
            "orc_build_h_dense": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _dp]),
            "orc_h_rows": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p, _i64p, _i32p,
                                          _dp]),
            "orc_eigh": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp]),
            "orc_lanczos": (ctypes.c_int, [ctypes.c_int64, _i64p, _i32p, _dp, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_int, _dp, _dp]),
            "orc_hb_select": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _dp, ctypes.c_double, _i64p, _i32p]),
            "orc_cmax": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp, _dp]),
            "orc_pt2": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                       ctypes.c_double, _dp, _i64p]),
            "orc_sample_rates": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, ctypes.c_int64, _i32p, ctypes.c_int64,
                                                ctypes.c_int, _dp]),
            "orc_fast_h_rows": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p, _i64p,
                                               _i32p, _dp]),
            "orc_fast_pt2": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _dp, ctypes.c_int, _dp, ctypes.c_int, _i32p,
                                            ctypes.c_int64, _dp, _i64p]),
            "orc_fast_sample": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _i32p, ctypes.c_int64, ctypes.c_int64,
                                               ctypes.c_int64, ctypes.c_int, _dp]),
            "orc_run": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                       ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _L = L
    return _L


def _p(a, t):
    return a.ctypes.data_as(t)


def set_threads(n):
    lib().orc_set_threads(int(n))


class OProblem:
    def __init__(self, path, restart=False):
        self.h = lib().orc_load(os.fsencode(path), int(restart))
        if not self.h:
            raise RuntimeError(lib().orc_error().decode())
        self.M = int(lib().orc_info_i(self.h, 0))
        self.neig = int(lib().orc_info_i(self.h, 1))
        self.n0 = int(lib().orc_info_i(self.h, 4))
        self.fcmax = int(lib().orc_info_i(self.h, 5))
        self.eps1 = float(lib().orc_info_d(self.h, 0))
        self.eps2 = float(lib().orc_info_d(self.h, 1))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)

    def fc_list(self, which):
        L = lib()
        out = []
        fc = ctypes.c_double()
        npow = ctypes.c_int()
        nsm = ctypes.c_int()
        fp = (ctypes.c_int * 6)()
        sm = (ctypes.c_int * 6)()
        mp = (ctypes.c_int * 6)()
        for i in range(L.orc_fc_count(self.h, which)):
            L.orc_fc_get(self.h, which, i, ctypes.byref(fc), ctypes.byref(npow), fp, ctypes.byref(nsm), sm, mp)
            out.append((fc.value, npow.value, [(sm[t], mp[t]) for t in range(nsm.value)]))
        return out

    def initial_basis(self):
        q = np.zeros((self.n0, self.M), dtype=np.int32)
        lib().orc_initial_basis(self.h, _p(q, _i32p))
        return q

    def build_h(self, q):
        q = np.ascontiguousarray(q, dtype=np.int32)
        n = q.shape[0]
        nnz = ctypes.c_int64()
        lib().orc_build_h(self.h, _p(q, _i32p), n, ctypes.byref(nnz), None, None, None)
        rp = np.zeros(n + 1, dtype=np.int64)
        col = np.zeros(nnz.value, dtype=np.int32)
        val = np.zeros(nnz.value)
        lib().orc_build_h(self.h, _p(q, _i32p), n, ctypes.byref(nnz), _p(rp, _i64p), _p(col, _i32p), _p(val, _dp))
        return rp, col, val

    def h_rows(self, q, rows):
        """rows of the all-pairs H over the whole basis q: (rowptr, col, val)"""
        q = np.ascontiguousarray(q, dtype=np.int32)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        nnz = ctypes.c_int64()
        if lib().orc_h_rows(self.h, _p(q, _i32p), q.shape[0], _p(rows, _i64p), len(rows), ctypes.byref(nnz), None,
                            None, None):
            raise RuntimeError(lib().orc_error().decode())
        rp = np.zeros(len(rows) + 1, dtype=np.int64)
        col = np.zeros(nnz.value, dtype=np.int32)
        val = np.zeros(nnz.value)
        lib().orc_h_rows(self.h, _p(q, _i32p), q.shape[0], _p(rows, _i64p), len(rows), ctypes.byref(nnz),
                         _p(rp, _i64p), _p(col, _i32p), _p(val, _dp))
        return rp, col, val

    def fast_h_rows(self, q, rows):
        """rows of H by the hash-based host form: (rowptr, col, val), no
        stored zeros"""
        q = np.ascontiguousarray(q, dtype=np.int32)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        nnz = ctypes.c_int64()
        if lib().orc_fast_h_rows(self.h, _p(q, _i32p), q.shape[0], _p(rows, _i64p), len(rows), ctypes.byref(nnz),
                                 None, None, None):
            raise RuntimeError(lib().orc_error().decode())
        rp = np.zeros(len(rows) + 1, dtype=np.int64)
        col = np.zeros(nnz.value, dtype=np.int32)
        val = np.zeros(nnz.value)
        lib().orc_fast_h_rows(self.h, _p(q, _i32p), q.shape[0], _p(rows, _i64p), len(rows), ctypes.byref(nnz),
                              _p(rp, _i64p), _p(col, _i32p), _p(val, _dp))
        return rp, col, val

    def fast_pt2(self, q, C, E, nopt, qA):
        """(PT2 corrections, pairs) over the external states qA by the
        hash-based host form"""
        q = np.ascontiguousarray(q, dtype=np.int32)
        C = np.ascontiguousarray(C, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        qA = np.ascontiguousarray(qA, dtype=np.int32)
        dE = np.zeros(nopt)
        npairs = ctypes.c_int64()
        if lib().orc_fast_pt2(self.h, _p(q, _i32p), q.shape[0], _p(C, _dp), C.shape[1], _p(E, _dp), nopt,
                              _p(qA, _i32p), qA.shape[0], _p(dE, _dp), ctypes.byref(npairs)):
            raise RuntimeError(lib().orc_error().decode())
        return dE, npairs.value

    def fast_sample(self, qV, qA, rows_h, rows_p, nopt):
        """timed sample of the hash-based host form (13 figures, see
        orc_fast_sample)"""
        qV = np.ascontiguousarray(qV, dtype=np.int32)
        qA = np.ascontiguousarray(qA, dtype=np.int32)
        out = np.zeros(13)
        if lib().orc_fast_sample(self.h, _p(qV, _i32p), qV.shape[0], _p(qA, _i32p), qA.shape[0], int(rows_h),
                                 int(rows_p), int(nopt), _p(out, _dp)):
            raise RuntimeError(lib().orc_error().decode())
        return out

    def build_h_dense(self, q):
        q = np.ascontiguousarray(q, dtype=np.int32)
        n = q.shape[0]
        H = np.zeros((n, n))
        lib().orc_build_h_dense(self.h, _p(q, _i32p), n, _p(H, _dp))
        return H

    def anharm_pot(self, which, idx, qn, qm):
        qn = np.ascontiguousarray(qn, dtype=np.int32)
        qm = np.ascontiguousarray(qm, dtype=np.int32)
        return lib().orc_anharm_pot(self.h, which, idx, _p(qn, _i32p), _p(qm, _i32p))

    def hb_select(self, q, cmax, eps):
        q = np.ascontiguousarray(q, dtype=np.int32)
        cmax = np.ascontiguousarray(cmax, dtype=np.float64)
        nn = ctypes.c_int64()
        if lib().orc_hb_select(self.h, _p(q, _i32p), q.shape[0], _p(cmax, _dp), eps, ctypes.byref(nn), None):
            raise RuntimeError(lib().orc_error().decode())
        out = np.zeros((nn.value, self.M), dtype=np.int32)
        lib().orc_hb_select(self.h, _p(q, _i32p), q.shape[0], _p(cmax, _dp), eps, ctypes.byref(nn),
                            _p(out, _i32p))
        return out

    def sample_rates(self, qV, rows, qA, nopt):
        """(H pairs/s, PT2 pairs/s, details) on a bounded sample."""
        qV = np.ascontiguousarray(qV, dtype=np.int32)
        qA = np.ascontiguousarray(qA, dtype=np.int32)
        out = np.zeros(4)
        lib().orc_sample_rates(self.h, _p(qV, _i32p), qV.shape[0], rows, _p(qA, _i32p), qA.shape[0], nopt,
                               _p(out, _dp))
        return out[0] / out[1], out[2] / out[3], out

    def pt2(self, q, C, E, nopt, eps2):
        q = np.ascontiguousarray(q, dtype=np.int32)
        C = np.ascontiguousarray(C, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        dE = np.zeros(nopt)
        ns2 = ctypes.c_int64()
        if lib().orc_pt2(self.h, _p(q, _i32p), q.shape[0], _p(C, _dp), C.shape[1], _p(E, _dp), nopt, eps2,
                         _p(dE, _dp), ctypes.byref(ns2)):
            raise RuntimeError(lib().orc_error().decode())
        return dE, ns2.value


def coeffsum(k, qn, qm):
    return lib().orc_coeffsum(k, qn, qm)


def eigh(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    Z = np.zeros((n, n))
    lib().orc_eigh(n, _p(A, _dp), _p(w, _dp), _p(Z, _dp))
    return w, Z


def lanczos(rp, col, val, nev, tol=1e-10, maxit=1000):
    n = rp.shape[0] - 1
    E = np.zeros(nev)
    C = np.zeros((n, nev))
    rc = lib().orc_lanczos(n, _p(rp, _i64p), _p(col, _i32p), _p(val, _dp), nev, tol, maxit, _p(E, _dp), _p(C, _dp))
    if rc:
        raise RuntimeError(lib().orc_error().decode())
    return E, C


def cmax(C, nopt):
    C = np.ascontiguousarray(C, dtype=np.float64)
    out = np.zeros(C.shape[0])
    lib().orc_cmax(C.shape[0], C.shape[1], nopt, _p(C, _dp), _p(out, _dp))
    return out


def run(inp, freqs, ckpt=None, threads=0, json_out=None, basis_out=None, quiet=True):
    rc = lib().orc_run(os.fsencode(inp), os.fsencode(freqs), os.fsencode(ckpt) if ckpt else None, int(threads),
                       os.fsencode(json_out) if json_out else None, os.fsencode(basis_out) if basis_out else None,
                       int(quiet))
    if rc:
        raise RuntimeError(lib().orc_error().decode())
