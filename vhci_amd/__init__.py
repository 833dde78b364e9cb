"""vhci_amd -- MI355X heat-bath VCI engine.

Python binding (ctypes) of ``libvhci.so``, the C ABI declared in
``include/vhci.h``.  The library runs every stage of the select -> build ->
diagonalize -> PT2 loop on the GPU; there is no host fallback: loading fails
loudly when the library has not been built, and every call raises when the
device reports an error.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvhci.so")
CLI_PATH = os.path.join(_HERE, "bin", "vhci")

FP64 = 0
MIXED = 1

_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); mirrors include/vhci.h
_SIGS = {
    "vhci_problem_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_vp)]),
    "vhci_problem_parse": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_vp)]),
    "vhci_problem_free": (None, [_vp]),
    "vhci_problem_error": (ctypes.c_char_p, []),
    "vhci_problem_int": (ctypes.c_int64, [_vp, ctypes.c_int]),
    "vhci_problem_real": (ctypes.c_double, [_vp, ctypes.c_int]),
    "vhci_problem_freqs": (ctypes.c_int, [_vp, _dp]),
    "vhci_problem_initial_basis": (ctypes.c_int, [_vp, _i32p]),
    "vhci_checkpoint_write": (ctypes.c_int, [_vp, ctypes.c_char_p, _i32p, ctypes.c_int64, _dp, ctypes.c_int, _dp,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "vhci_checkpoint_read": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                            _i32p, ctypes.c_int64, _dp, _dp, ctypes.c_int64]),
    "vhci_problem_fc_count": (ctypes.c_int, [_vp, ctypes.c_int]),
    "vhci_problem_fc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _ip, _ip, _ip, _ip]),
    "vhci_ladder_sum": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "vhci_ctx_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "vhci_ctx_destroy": (None, [_vp]),
    "vhci_last_error": (ctypes.c_char_p, [_vp]),
    "vhci_set_basis": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64]),
    "vhci_basis_size": (ctypes.c_int64, [_vp]),
    "vhci_get_basis": (ctypes.c_int, [_vp, _i32p]),
    "vhci_build_h": (ctypes.c_int, [_vp, _i64p]),
    "vhci_matrix_free": (ctypes.c_int, [_vp]),
    "vhci_structural_nnz": (ctypes.c_int64, [_vp]),
    "vhci_get_h": (ctypes.c_int, [_vp, _i64p, _i64p, _i32p, _dp, _dp]),
    "vhci_get_h_rows": (ctypes.c_int, [_vp, _i64p, ctypes.c_int64, _i64p, _i64p, _i32p, _dp, _dp]),
    "vhci_eigsolve": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp]),
    "vhci_dense_diag": (ctypes.c_int, [_vp, _dp, _dp]),
    "vhci_ritz_cols": (ctypes.c_int, [_vp]),
    "vhci_get_vectors": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "vhci_set_vectors": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "vhci_freq_assignments": (ctypes.c_int, [_vp, ctypes.c_int, _i64p, _i64p, _dp, ctypes.c_int64, _i64p]),
    "vhci_hb_select": (ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_double, _i64p]),
    "vhci_get_new_states": (ctypes.c_int, [_vp, _i32p]),
    "vhci_append_new_states": (ctypes.c_int, [_vp]),
    "vhci_pt2": (ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_double, _dp, _i64p]),
    "vhci_apply_h": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "vhci_apply_h_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "vhci_time_spmm": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "vhci_copy_probe": (ctypes.c_int, [_vp, ctypes.c_int64, _dp]),
    "vhci_spmm_profile": (ctypes.c_int, [_vp, ctypes.c_int]),
    "vhci_spmm_profile_read": (ctypes.c_int, [_vp, _dp]),
    "vhci_gemm_profile_read": (ctypes.c_int, [_vp, _dp]),
    "vhci_phase_stats": (ctypes.c_int, [_vp, _dp]),
    "vhci_pool_stats": (ctypes.c_int, [_vp, _i64p]),
    "vhci_solve": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _ip, _i64p, _i64p, _dp]),
    "vhci_run_ctx": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _dp, _dp, _ip, _i64p, _dp]),
    "vhci_iteration_sizes": (ctypes.c_int, [_vp, _i64p, ctypes.c_int]),
    "vhci_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "vhci_server_destroy": (None, [_vp]),
    "vhci_server_error": (ctypes.c_char_p, [_vp]),
    "vhci_server_submit": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_double, _i64p]),
    "vhci_server_wait": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_double, _ip]),
    "vhci_server_result": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp, _ip, _i64p, _dp]),
    "vhci_server_stats": (ctypes.c_int, [_vp, _dp]),
    "vhci_main": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]),
}

EXPORTED = tuple(_SIGS)


class VhciError(RuntimeError):
    pass


def lib():
    """Load libvhci.so (built in-tree); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VhciError("libvhci.so is not built; run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Problem:
    """A parsed input file (the reference's ReadCIInput globals)."""

    def __init__(self, path=None, text=None, restart=False):
        L = lib()
        h = _vp()
        if text is not None:
            rc = L.vhci_problem_parse(text.encode(), int(restart), ctypes.byref(h))
        else:
            rc = L.vhci_problem_load(os.fsencode(path), int(restart), ctypes.byref(h))
        if rc != 0:
            raise VhciError(L.vhci_problem_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.vhci_problem_free(self.h)
            self.h = None

    def _i(self, k):
        return int(lib().vhci_problem_int(self.h, k))

    M = property(lambda s: s._i(0))
    neig = property(lambda s: s._i(1))
    ntot = property(lambda s: s._i(2))
    perturb = property(lambda s: bool(s._i(3)))
    n0 = property(lambda s: s._i(4))
    fcmax = property(lambda s: s._i(5))
    nfc = property(lambda s: s._i(6))
    nhb = property(lambda s: s._i(7))
    npatterns = property(lambda s: s._i(8))
    eps1 = property(lambda s: float(lib().vhci_problem_real(s.h, 0)))
    eps2 = property(lambda s: float(lib().vhci_problem_real(s.h, 1)))

    def freqs(self):
        f = np.zeros(self.M)
        lib().vhci_problem_freqs(self.h, _ptr(f, _dp))
        return f

    def initial_basis(self):
        q = np.zeros((self.n0, self.M), dtype=np.int32)
        lib().vhci_problem_initial_basis(self.h, _ptr(q, _i32p))
        return q

    def write_checkpoint(self, path, q, E, C, threads=1, binary=False):
        """The checkpoint file of basis q (n x M), energies E and coefficients
        C (n x ncol), as the reference's WriteCheckpoint writes it (or the
        opt-in binary form)."""
        q = np.ascontiguousarray(q, dtype=np.int32)
        E = np.ascontiguousarray(E, dtype=np.float64)
        C = np.ascontiguousarray(C, dtype=np.float64)
        if C.ndim != 2 or q.shape != (C.shape[0], self.M):
            raise ValueError("q must be n x M and C n x ncol")
        rc = lib().vhci_checkpoint_write(self.h, os.fsencode(path), _ptr(q, _i32p), q.shape[0], _ptr(E, _dp), E.size,
                                         _ptr(C, _dp), C.shape[1], int(threads), int(bool(binary)))
        if rc != 0:
            raise VhciError(lib().vhci_problem_error().decode())

    def read_checkpoint(self, path, threads=1):
        """(q n x M, E NEig, C n x NEig) from a checkpoint file of either form."""
        L = lib()
        n = ctypes.c_int64()
        rc = L.vhci_checkpoint_read(self.h, os.fsencode(path), int(threads), ctypes.byref(n), None, 0, None, None, 0)
        if rc == -2:
            raise VhciError(L.vhci_problem_error().decode())
        if n.value < 1:
            raise VhciError("Error: checkpoint file is truncated.")
        q = np.zeros((n.value, self.M), dtype=np.int32)
        E = np.zeros(self.neig)
        C = np.zeros((n.value, self.neig))
        rc = L.vhci_checkpoint_read(self.h, os.fsencode(path), int(threads), ctypes.byref(n), _ptr(q, _i32p), q.size,
                                    _ptr(E, _dp), _ptr(C, _dp), C.size)
        if rc != 0:
            raise VhciError(L.vhci_problem_error().decode())
        return q, E, C

    def fc_list(self, which):
        """[(value, order, [(mode, power), ...]), ...] of one FC list."""
        L = lib()
        n = L.vhci_problem_fc_count(self.h, which)
        out = []
        fc = ctypes.c_double()
        order = ctypes.c_int()
        nsm = ctypes.c_int()
        sm = (ctypes.c_int * 6)()
        mp = (ctypes.c_int * 6)()
        for i in range(n):
            L.vhci_problem_fc(self.h, which, i, ctypes.byref(fc), ctypes.byref(order), ctypes.byref(nsm), sm, mp)
            out.append((fc.value, order.value, [(sm[t], mp[t]) for t in range(nsm.value)]))
        return out


def ladder_sum(k, qs, qe):
    return float(lib().vhci_ladder_sum(k, qs, qe))


class Context:
    """Device state of one solve (basis hash, H, Ritz block) on one GPU."""

    def __init__(self, problem, form=FP64, device=0):
        L = lib()
        h = _vp()
        rc = L.vhci_ctx_create(problem.h, form, device, ctypes.byref(h))
        if rc != 0:
            raise VhciError(L.vhci_problem_error().decode())
        self.h = h
        self.problem = problem
        self.M = problem.M

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.vhci_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc):
        if rc != 0:
            raise VhciError("%s (code %d)" % (lib().vhci_last_error(self.h).decode(), rc))

    @property
    def n(self):
        return int(lib().vhci_basis_size(self.h))

    def set_basis(self, q):
        q = np.ascontiguousarray(q, dtype=np.int32)
        self._check(lib().vhci_set_basis(self.h, _ptr(q, _i32p), q.shape[0]))

    def basis(self):
        q = np.zeros((self.n, self.M), dtype=np.int32)
        self._check(lib().vhci_get_basis(self.h, _ptr(q, _i32p)))
        return q

    def build_h(self):
        st = np.zeros(4, dtype=np.int64)
        self._check(lib().vhci_build_h(self.h, _ptr(st, _i64p)))
        return {"n": int(st[0]), "nnz_offdiag": int(st[1]), "nnz_struct": int(st[2]), "us": int(st[3])}

    @property
    def matrix_free(self):
        """True when the built H keeps no stored entries (products regenerate them)."""
        r = int(lib().vhci_matrix_free(self.h))
        if r < 0:
            self._check(r)
        return bool(r)

    def structural_nnz(self):
        r = int(lib().vhci_structural_nnz(self.h))
        if r < -1:
            self._check(r)
        return r

    def get_h(self):
        L = lib()
        nnz = ctypes.c_int64()
        self._check(L.vhci_get_h(self.h, ctypes.byref(nnz), None, None, None, None))
        n = self.n
        rp = np.zeros(n + 1, dtype=np.int64)
        col = np.zeros(max(1, nnz.value), dtype=np.int32)
        val = np.zeros(max(1, nnz.value))
        diag = np.zeros(n)
        self._check(L.vhci_get_h(self.h, ctypes.byref(nnz), _ptr(rp, _i64p), _ptr(col, _i32p), _ptr(val, _dp),
                                 _ptr(diag, _dp)))
        return rp, col[: nnz.value], val[: nnz.value], diag

    def get_h_rows(self, rows):
        """(rowptr, col, val, diag) of the given basis rows (off-diagonal
        entries, columns in basis order)"""
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        nnz = ctypes.c_int64()
        self._check(lib().vhci_get_h_rows(self.h, _ptr(rows, _i64p), len(rows), ctypes.byref(nnz), None, None, None,
                                          None))
        rp = np.zeros(len(rows) + 1, dtype=np.int64)
        col = np.zeros(nnz.value, dtype=np.int32)
        val = np.zeros(nnz.value)
        diag = np.zeros(len(rows))
        self._check(lib().vhci_get_h_rows(self.h, _ptr(rows, _i64p), len(rows), ctypes.byref(nnz), _ptr(rp, _i64p),
                                          _ptr(col, _i32p), _ptr(val, _dp), _ptr(diag, _dp)))
        return rp, col, val, diag

    def apply_h(self, X, fp32=False):
        """H X for an n x b block (or an n-vector) in basis order, through
        the solver's product kernels (fp32: the mixed form's FP32 product)."""
        X = np.asarray(X, dtype=np.float64)
        vec = X.ndim == 1
        X = np.ascontiguousarray(X.reshape(self.n, -1))
        Y = np.zeros_like(X)
        self._check(lib().vhci_apply_h(self.h, X.shape[1], int(fp32), _ptr(X, _dp), _ptr(Y, _dp)))
        return Y[:, 0] if vec else Y

    def apply_h_noalloc(self, X, fp32=False):
        """apply_h through the entry point that allocates no device memory
        (vhci_apply_h_noalloc)"""
        X = np.asarray(X, dtype=np.float64)
        vec = X.ndim == 1
        X = np.ascontiguousarray(X.reshape(self.n, -1))
        Y = np.zeros_like(X)
        self._check(lib().vhci_apply_h_noalloc(self.h, X.shape[1], int(fp32), _ptr(X, _dp), _ptr(Y, _dp)))
        return Y[:, 0] if vec else Y

    def eigsolve(self, nev, tol=1e-10, maxit=500, vectors=True):
        n = self.n
        nev = min(nev, n)
        E = np.zeros(nev)
        C = np.zeros((n, nev)) if vectors else None
        self._check(lib().vhci_eigsolve(self.h, nev, tol, maxit, _ptr(E, _dp), _ptr(C, _dp) if vectors else None))
        return E, C

    def dense_diag(self, vectors=True):
        n = self.n
        E = np.zeros(n)
        C = np.zeros((n, n)) if vectors else None
        self._check(lib().vhci_dense_diag(self.h, _ptr(E, _dp), _ptr(C, _dp) if vectors else None))
        return E, C

    def vectors(self, ncol=None):
        ncol = ncol or int(lib().vhci_ritz_cols(self.h))
        C = np.zeros((self.n, ncol))
        self._check(lib().vhci_get_vectors(self.h, _ptr(C, _dp), ncol))
        return C

    def freq_assignments(self, ncol):
        """PrintFreqs' assignment scan on the device: per eigenvector, the
        basis rows with |C| >= max |C| / 2 (ascending) and their |C|."""
        L = lib()
        counts = np.zeros(ncol, dtype=np.int64)
        total = np.zeros(1, dtype=np.int64)
        rc = L.vhci_freq_assignments(self.h, ncol, _ptr(counts, _i64p), None, None, 0, _ptr(total, _i64p))
        if rc not in (0, -1):
            self._check(rc)
        rows = np.zeros(max(1, int(total[0])), dtype=np.int64)
        vals = np.zeros(max(1, int(total[0])))
        self._check(L.vhci_freq_assignments(self.h, ncol, _ptr(counts, _i64p), _ptr(rows, _i64p), _ptr(vals, _dp),
                                            rows.size, _ptr(total, _i64p)))
        off = np.concatenate([[0], np.cumsum(counts)])
        return [(rows[off[c]:off[c + 1]].copy(), vals[off[c]:off[c + 1]].copy()) for c in range(ncol)]

    def set_vectors(self, C):
        C = np.ascontiguousarray(C, dtype=np.float64)
        self._check(lib().vhci_set_vectors(self.h, _ptr(C, _dp), C.shape[1]))

    def hb_select(self, eps, nopt=None, cmax=None):
        nnew = ctypes.c_int64()
        if cmax is not None:
            cmax = np.ascontiguousarray(cmax, dtype=np.float64)
        self._check(lib().vhci_hb_select(self.h, _ptr(cmax, _dp) if cmax is not None else None, int(nopt or 0),
                                         float(eps), ctypes.byref(nnew)))
        return nnew.value

    def new_states(self, count):
        q = np.zeros((count, self.M), dtype=np.int32)
        self._check(lib().vhci_get_new_states(self.h, _ptr(q, _i32p)))
        return q

    def append_new_states(self):
        self._check(lib().vhci_append_new_states(self.h))

    def pt2(self, E, nopt, eps2):
        E = np.ascontiguousarray(E[:nopt], dtype=np.float64)
        dE = np.zeros(nopt)
        ns2 = ctypes.c_int64()
        self._check(lib().vhci_pt2(self.h, _ptr(E, _dp), nopt, float(eps2), _ptr(dE, _dp), ctypes.byref(ns2)))
        return dE, ns2.value

    def run(self, tol=1e-10, restart=False):
        """The whole select -> build -> diagonalize -> PT2 loop on this context."""
        nmax = max(self.problem.neig, self.n, 1)
        E = np.zeros(nmax)
        Ev = np.zeros(nmax)
        nE = ctypes.c_int()
        info = np.zeros(4, dtype=np.int64)
        t = np.zeros(8)
        self._check(lib().vhci_run_ctx(self.h, tol, int(restart), _ptr(E, _dp), _ptr(Ev, _dp), ctypes.byref(nE),
                                       _ptr(info, _i64p), _ptr(t, _dp)))
        keys = ["total_ms", "build_ms", "eig_ms", "hb_ms", "pt2_select_ms", "pt2_sum_ms", "iterations",
                "spmm_vectors"]
        return {"E": E[: nE.value].copy(), "E_var": Ev[: nE.value].copy(), "nvar": int(info[0]),
                "ns2": int(info[1]), "hb_iterations": int(info[2]), "checksum": int(info[3]) & ((1 << 64) - 1),
                "timings": dict(zip(keys, t.tolist()))}

    def iteration_sizes(self):
        out = np.zeros(64, dtype=np.int64)
        k = lib().vhci_iteration_sizes(self.h, _ptr(out, _i64p), 64)
        return [int(x) for x in out[:k]]

    def time_spmm(self, b, reps=20, fp32=False):
        """(ms per product, algorithmic bytes per product, launches per product)"""
        out = np.zeros(3)
        self._check(lib().vhci_time_spmm(self.h, b, reps, int(fp32), _ptr(out, _dp)))
        return float(out[0]), float(out[1]), int(out[2])

    def spmm_profile(self, b):
        """Start (b > 0, clears) or stop (b = 0) per-launch timing of the
        solver's b-column SpMMs."""
        self._check(lib().vhci_spmm_profile(self.h, int(b)))

    def spmm_profile_read(self):
        """(total ms, total algorithmic bytes, launches) since the start"""
        out = np.zeros(3)
        self._check(lib().vhci_spmm_profile_read(self.h, _ptr(out, _dp)))
        return float(out[0]), float(out[1]), int(out[2])

    def gemm_profile_read(self):
        """(total ms, flops, calls) of the Ritz rotation GEMMs while SpMM
        profiling is on"""
        out = np.zeros(3)
        self._check(lib().vhci_gemm_profile_read(self.h, _ptr(out, _dp)))
        return float(out[0]), float(out[1]), int(out[2])

    def phase_stats(self):
        out = np.zeros(16)
        self._check(lib().vhci_phase_stats(self.h, _ptr(out, _dp)))
        keys = ["n", "nnz", "patterns", "build_ms", "hb_candidates", "hb_unique", "hb_ms", "ns2", "pairs",
                "pt2_sum_ms", "key_bytes", "nstate", "probed_rows", "carried", "pt2_chunks", "h_storage"]
        return dict(zip(keys, out.tolist()))

    def pool_stats(self):
        """device-memory pool of the context: hipMalloc calls, hipFree calls,
        idle blocks handed out again, bytes held, idle bytes"""
        out = np.zeros(5, dtype=np.int64)
        self._check(lib().vhci_pool_stats(self.h, _ptr(out, _i64p)))
        return dict(zip(["mallocs", "frees", "reuses", "held", "idle"], out.tolist()))

    def copy_probe(self, nbytes):
        ms = ctypes.c_double()
        self._check(lib().vhci_copy_probe(self.h, int(nbytes), ctypes.byref(ms)))
        return ms.value


def basis_checksum(q):
    """Checksum of a basis (int32 rows) as vhci_run_ctx reports it."""
    v = np.ascontiguousarray(q, dtype=np.int32).ravel().astype(np.int64).astype(np.uint64)
    with np.errstate(over="ignore"):
        pw = np.cumprod(np.full(v.size, 1099511628211, dtype=np.uint64), dtype=np.uint64)
        return int(np.sum(v * pw, dtype=np.uint64))


def solve(problem, form=FP64, device=0, tol=1e-10):
    """Whole select -> build -> diagonalize -> PT2 run of one problem."""
    L = lib()
    nmax = max(problem.neig, problem.n0, 1)
    E = np.zeros(nmax)
    nE = ctypes.c_int()
    nvar = ctypes.c_int64()
    ns2 = ctypes.c_int64()
    t = np.zeros(8)
    rc = L.vhci_solve(problem.h, form, device, tol, _ptr(E, _dp), ctypes.byref(nE), ctypes.byref(nvar),
                      ctypes.byref(ns2), _ptr(t, _dp))
    if rc != 0:
        raise VhciError(L.vhci_problem_error().decode())
    keys = ["total_ms", "build_ms", "eig_ms", "hb_ms", "pt2_select_ms", "pt2_sum_ms", "iterations", "spmm_vectors"]
    return {"E": E[: nE.value].copy(), "nvar": nvar.value, "ns2": ns2.value, "timings": dict(zip(keys, t.tolist()))}


class Server:
    """The serving loop: requests (input text, form, tolerance) queue FIFO
    and run on `slots` worker threads, each request on its own device
    context and stream, admitted when a slot is free and its predicted device
    bytes fit under mem_limit next to the running requests' reservations."""

    STATES = ("queued", "running", "done", "failed")

    def __init__(self, device=0, slots=4, mem_limit=200 << 30):
        L = lib()
        h = _vp()
        rc = L.vhci_server_create(device, slots, int(mem_limit), ctypes.byref(h))
        if rc != 0:
            raise VhciError(L.vhci_problem_error().decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().vhci_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise VhciError(lib().vhci_server_error(self.h).decode())

    def submit(self, text, form=FP64, tol=0.0):
        rid = ctypes.c_int64()
        self._check(lib().vhci_server_submit(self.h, text.encode(), form, tol, ctypes.byref(rid)))
        return rid.value

    def wait(self, rid, timeout=-1.0):
        done = ctypes.c_int()
        self._check(lib().vhci_server_wait(self.h, rid, float(timeout), ctypes.byref(done)))
        return bool(done.value)

    def result(self, rid):
        """Record of a request; raises VhciError if it failed."""
        E = np.zeros(1024)
        Ev = np.zeros(1024)
        nE = ctypes.c_int()
        info = np.zeros(6, dtype=np.int64)
        t = np.zeros(6)
        rc = lib().vhci_server_result(self.h, rid, _ptr(E, _dp), _ptr(Ev, _dp), ctypes.byref(nE), _ptr(info, _i64p),
                                      _ptr(t, _dp))
        self._check(rc)
        return {"state": self.STATES[int(info[0])], "E": E[: nE.value].copy(), "E_var": Ev[: nE.value].copy(),
                "nvar": int(info[1]), "ns2": int(info[2]), "hb_iterations": int(info[3]),
                "checksum": int(info[4]) & ((1 << 64) - 1), "slot": int(info[5]), "t_submit": float(t[0]),
                "t_start": float(t[1]), "t_end": float(t[2]), "predicted_bytes": float(t[3]),
                "peak_bytes": float(t[4]), "reserved_bytes": float(t[5])}

    def stats(self):
        out = np.zeros(5)
        self._check(lib().vhci_server_stats(self.h, _ptr(out, _dp)))
        return {"running": int(out[0]), "queued": int(out[1]), "reserved_bytes": float(out[2]),
                "max_running": int(out[3]), "finished": int(out[4])}

