"""ctypes wrapper of liboracle.so (the FP64 CPU oracle) and oracle/_ref/libref.so.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_zscore.argtypes = [f32p, C.c_int64, C.c_int, f64p, f32p]
        L.orc_r.argtypes = [f64p, C.c_int, C.c_int64, C.c_int64]
        L.orc_r.restype = C.c_double
        L.orc_r_threshold_eff.argtypes = [C.c_float, C.c_int]
        L.orc_r_threshold_eff.restype = C.c_float
        L.orc_sparsity_k.argtypes = [C.c_int64, C.c_float]
        L.orc_sparsity_k.restype = C.c_int64
        L.orc_corr_count_r.argtypes = [f64p, C.c_int64, C.c_int, C.c_float]
        L.orc_corr_count_r.restype = C.c_int64
        L.orc_corr_count_rows.argtypes = [f64p, C.c_int64, C.c_int, C.c_float, C.c_int64, C.c_int64]
        L.orc_corr_count_rows.restype = C.c_int64
        L.orc_corr_rows_r.argtypes = [f64p, C.c_int64, C.c_int, C.c_float, i32p, C.c_int64, i32p, i32p,
                                      C.c_void_p]
        L.orc_cp_weighted_at.argtypes = [i32p, i32p, f32p, C.c_int64, C.c_int, C.c_int, i32p, C.c_int64, f64p]
        L.orc_corr_csr_r.argtypes = [f64p, C.c_int64, C.c_int, C.c_float, i32p, i32p, C.c_void_p]
        L.orc_corr_csr_s.argtypes = [f64p, C.c_int64, C.c_int, C.c_int64, i32p, i32p, C.c_void_p]
        L.orc_corr_csr_s.restype = C.c_double
        L.orc_degree.argtypes = [i32p, C.c_int64, i32p, f32p]
        L.orc_strength.argtypes = [i32p, f32p, C.c_int64, f32p]
        L.orc_triangles.argtypes = [i32p, i32p, C.c_int64, i64p]
        L.orc_cp_binary.argtypes = [i32p, i32p, C.c_int64, f32p]
        L.orc_cp_binary.restype = C.c_double
        L.orc_cp_weighted.argtypes = [i32p, i32p, f32p, C.c_int64, C.c_int, C.c_int, f64p]
        L.orc_cp_weighted.restype = C.c_double
        L.orc_bc.argtypes = [i32p, i32p, C.c_int64, C.c_void_p, C.c_int64, f64p]
        L.orc_bfs.argtypes = [i32p, i32p, C.c_int64, C.c_int32, i32p]
        L.orc_bfs_efficiency.argtypes = [i32p, i32p, C.c_int64, f32p, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32)]
        L.orc_bfs_efficiency.restype = C.c_double
        L.orc_fw_blocked.argtypes = [f32p, C.c_int64, C.c_int64]
        L.orc_fw_init.argtypes = [i32p, i32p, C.c_void_p, C.c_int64, C.c_int, f32p]
        L.orc_modularity.argtypes = [i32p, i32p, C.c_void_p, C.c_int64, C.c_int, i32p,
                                     C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_int32)]
        L.orc_modularity.restype = C.c_double
        L.orc_modularity_ex.argtypes = [i32p, i32p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_int,
                                        C.c_int, i32p, C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_int32)]
        L.orc_modularity_ex.restype = C.c_double
        L.orc_modularity_q.argtypes = [i32p, i32p, C.c_void_p, C.c_int64, C.c_int, i32p]
        L.orc_modularity_q.restype = C.c_double
        L.orc_ec.argtypes = [i32p, i32p, C.c_void_p, C.c_int64, f64p, f64p, C.POINTER(C.c_int32)]
        L.orc_ec.restype = C.c_double
        L.orc_pc.argtypes = [i32p, i32p, C.c_void_p, C.c_int64, i32p, f32p, f32p]
        L.orc_pc.restype = C.c_int32
        L.orc_rewire.argtypes = [i32p, i32p, C.c_void_p, C.c_int64, C.c_uint64, C.c_int64, i32p, C.c_void_p]
        L.orc_rewire.restype = C.c_int64
        L.orc_cormat_add.argtypes = [f64p, C.c_int64, C.c_int, C.c_int, f32p]
        L.orc_cormat_finish.argtypes = [f32p, C.c_int64, C.c_int, C.c_int]
        L.orc_dense_csr_r.argtypes = [f32p, C.c_int64, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_dense_csr_r.restype = C.c_int64
        L.orc_dense_csr_s.argtypes = [f32p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_dense_csr_s.restype = C.c_float
        _LIB = L
    return _LIB


def ref():
    """The compiled reference CPU functions, or None when not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_bc_all.argtypes = [i32p, i32p, C.c_int32, f32p]
        R.ref_fw_dense.argtypes = [f32p, C.c_int32, C.c_int32]
        R.ref_fw_init.argtypes = [i32p, i32p, f32p, C.c_int32, C.c_int32, f32p]
        _REF = R
    return _REF


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ wrappers

def set_threads(n: int) -> None:
    """Threads of the oracle's OpenMP loops (0: all processors)."""
    lib().orc_set_threads(n)


def get_threads() -> int:
    return int(lib().orc_get_threads())


def zscore(bold_tn: np.ndarray):
    bold_tn = np.ascontiguousarray(bold_tn, dtype=np.float32)
    t, n = bold_tn.shape
    z64 = np.empty((n, t), np.float64)
    z32 = np.empty((n, t), np.float32)
    lib().orc_zscore(bold_tn, n, t, z64, z32)
    return z64, z32


def r_pair(z64: np.ndarray, i: int, j: int) -> float:
    return lib().orc_r(np.ascontiguousarray(z64), z64.shape[1], i, j)


def threshold_eff(thr: float, weighted: bool) -> float:
    return float(lib().orc_r_threshold_eff(thr, int(weighted)))


def sparsity_k(n: int, s_percent: float) -> int:
    return int(lib().orc_sparsity_k(n, s_percent))


def corr_count_rows(z64: np.ndarray, thr: float, i0: int, i1: int) -> int:
    z64 = np.ascontiguousarray(z64)
    n, t = z64.shape
    return int(lib().orc_corr_count_rows(z64, n, t, thr, i0, i1))


def corr_csr_r(z64: np.ndarray, thr: float, weighted: bool = False):
    """(row, col, val-or-None) for the r-threshold rule."""
    z64 = np.ascontiguousarray(z64)
    n, t = z64.shape
    th = threshold_eff(thr, weighted)
    nnz = lib().orc_corr_count_r(z64, n, t, th)
    row = np.empty(n + 1, np.int32)
    col = np.empty(max(nnz, 1), np.int32)
    val = np.empty(max(nnz, 1), np.float32) if weighted else None
    lib().orc_corr_csr_r(z64, n, t, th, row, col, _ptr(val))
    return row, col[:nnz], (val[:nnz] if weighted else None)


def corr_rows_r(z64: np.ndarray, thr: float, rows, weighted: bool = False):
    """The rows `rows` of the r-threshold CSR, each formed alone: a list of
    (cols, vals-or-None) per row."""
    z64 = np.ascontiguousarray(z64)
    n, t = z64.shape
    rows = np.ascontiguousarray(rows, np.int32)
    th = threshold_eff(thr, weighted)
    deg = np.empty(rows.size, np.int32)
    col = np.empty((rows.size, n), np.int32)
    val = np.empty((rows.size, n), np.float32) if weighted else None
    lib().orc_corr_rows_r(z64, n, t, th, rows, rows.size, deg, col, _ptr(val))
    return [(col[q, :deg[q]].copy(), None if val is None else val[q, :deg[q]].copy()) for q in range(rows.size)]


def corr_csr_s(z64: np.ndarray, s_percent: float, weighted: bool = False):
    """(row, col, val-or-None, t_star) for the sparsity rule."""
    z64 = np.ascontiguousarray(z64)
    n, t = z64.shape
    k = sparsity_k(n, s_percent)
    row = np.empty(n + 1, np.int32)
    col = np.empty(max(2 * k, 1), np.int32)
    val = np.empty(max(2 * k, 1), np.float32) if weighted else None
    tstar = lib().orc_corr_csr_s(z64, n, t, k, row, col, _ptr(val))
    nnz = int(row[-1])
    return row, col[:nnz], (val[:nnz] if weighted else None), float(tstar)


def degree(row: np.ndarray):
    n = row.size - 1
    d = np.empty(n, np.int32)
    dr = np.empty(n, np.float32)
    lib().orc_degree(np.ascontiguousarray(row, np.int32), n, d, dr)
    return d, dr


def strength(row, val):
    n = row.size - 1
    s = np.empty(n, np.float32)
    lib().orc_strength(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(val, np.float32), n, s)
    return s


def triangles(row, col):
    n = row.size - 1
    t = np.empty(n, np.int64)
    lib().orc_triangles(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), n, t)
    return t


def cp_binary(row, col):
    n = row.size - 1
    cp = np.empty(n, np.float32)
    mean = lib().orc_cp_binary(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), n, cp)
    return cp, mean


def cp_weighted(row, col, val, kind=2, wmax_mode=0):
    n = row.size - 1
    cp = np.empty(n, np.float64)
    mean = lib().orc_cp_weighted(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                                 np.ascontiguousarray(val, np.float32), n, kind, wmax_mode, cp)
    return cp, mean


def cp_weighted_at(row, col, val, verts, kind=2, wmax_mode=0):
    """cp_weighted at the vertices verts only."""
    n = row.size - 1
    verts = np.ascontiguousarray(verts, np.int32)
    cp = np.empty(verts.size, np.float64)
    lib().orc_cp_weighted_at(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                             np.ascontiguousarray(val, np.float32), n, kind, wmax_mode, verts, verts.size, cp)
    return cp


def bc(row, col, sources=None):
    n = row.size - 1
    out = np.empty(n, np.float64)
    src = None if sources is None else np.ascontiguousarray(sources, np.int32)
    lib().orc_bc(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), n,
                 _ptr(src), 0 if src is None else src.size, out)
    return out


def bfs(row, col, src):
    n = row.size - 1
    lev = np.empty(n, np.int32)
    lib().orc_bfs(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), n, src, lev)
    return lev


def bfs_efficiency(row, col):
    """(eff float32[n], Lp, isolated, unconnected) of a binary network."""
    n = row.size - 1
    eff = np.empty(n, np.float32)
    iso, unc = C.c_int32(0), C.c_int32(0)
    lp = lib().orc_bfs_efficiency(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), n,
                                  eff, C.byref(iso), C.byref(unc))
    return eff, lp, iso.value, unc.value


def fw_init(row, col, val, kind):
    n = row.size - 1
    d = np.empty((n, n), np.float32)
    v = None if val is None else np.ascontiguousarray(val, np.float32)
    lib().orc_fw_init(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), _ptr(v), n, kind, d)
    return d


def fw_blocked(d: np.ndarray, b: int) -> np.ndarray:
    d = np.array(d, dtype=np.float32, order="C", copy=True)
    lib().orc_fw_blocked(d, d.shape[0], b)
    return d


def modularity(row, col, val=None, kmode=0, stop=0, tol=1e-8, max_iter=10000, max_depth=0):
    """(labels, Q, lambda0, round-0 vector, rounds); stop 0 = the reference's
    vector rule, 1 = the Rayleigh-quotient rule with relative tolerance tol;
    max_depth > 0 stops the bisection that many levels down."""
    n = row.size - 1
    lab = np.empty(n, np.int32)
    lam = C.c_double(0.0)
    nr = C.c_int32(0)
    eig = np.zeros(n, np.float64)
    v = None if val is None else np.ascontiguousarray(val, np.float32)
    q = lib().orc_modularity_ex(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                                _ptr(v), n, kmode, stop, tol, max_iter, max_depth, lab, C.byref(lam), _ptr(eig),
                                C.byref(nr))
    return lab, q, lam.value, eig, nr.value


def modularity_q(row, col, labels, val=None, kmode=0):
    n = row.size - 1
    v = None if val is None else np.ascontiguousarray(val, np.float32)
    return lib().orc_modularity_q(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                                  _ptr(v), n, kmode, np.ascontiguousarray(labels, np.int32))



def ec(row, col, val=None):
    """(vhat, avec, eigenvalue, iterations) of the reference's power iteration."""
    n = row.size - 1
    vh, av = np.empty(n, np.float64), np.empty(n, np.float64)
    it = C.c_int32(0)
    v = None if val is None else np.ascontiguousarray(val, np.float32)
    lam = lib().orc_ec(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), _ptr(v), n,
                       vh, av, C.byref(it))
    return vh, av, lam, it.value


def pc(row, col, labels, val=None):
    """(pc, pc_normalized) float32 and the module count K."""
    n = row.size - 1
    p, q = np.empty(n, np.float32), np.empty(n, np.float32)
    v = None if val is None else np.ascontiguousarray(val, np.float32)
    k = lib().orc_pc(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), _ptr(v), n,
                     np.ascontiguousarray(labels, np.int32), p, q)
    return p, q, k


def rewire(row, col, val=None, seed=0, attempts=0):
    """Sequential degree-preserving rewiring: (col, val or None, swaps)."""
    n = row.size - 1
    oc = np.empty(col.size, np.int32)
    v = None if val is None else np.ascontiguousarray(val, np.float32)
    ov = None if val is None else np.empty(col.size, np.float32)
    sw = lib().orc_rewire(np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32), _ptr(v), n,
                          seed & (2**64 - 1), attempts, oc, _ptr(ov))
    return oc, ov, sw


def cormat_average(subjects, fisher=False):
    """Packed upper-triangle float32 average over subjects (list of [T, N]
    time-major float32 series with the same N)."""
    n = subjects[0].shape[1]
    acc = np.zeros(n * (n - 1) // 2, np.float32)
    for x in subjects:
        z64, _ = zscore(x)
        lib().orc_cormat_add(np.ascontiguousarray(z64), n, x.shape[0], int(fisher), acc)
    lib().orc_cormat_finish(acc, acc.size, len(subjects), int(fisher))
    return acc


def dense_csr_r(a, n, thr, weighted=False):
    """(row, col, val-or-None) of a > thr on a packed float32 matrix."""
    a = np.ascontiguousarray(a, np.float32)
    nnz = lib().orc_dense_csr_r(a, n, thr, None, None, None)
    row, col = np.empty(n + 1, np.int32), np.empty(max(nnz, 1), np.int32)
    val = np.empty(max(nnz, 1), np.float32) if weighted else None
    lib().orc_dense_csr_r(a, n, thr, _ptr(row), _ptr(col), _ptr(val))
    return row, col[:nnz], (val[:nnz] if weighted else None)


def dense_csr_s(a, n, s_percent, weighted=False):
    """(row, col, val-or-None, t) keeping the k largest, k from the sparsity rule."""
    a = np.ascontiguousarray(a, np.float32)
    k = sparsity_k(n, s_percent)
    row, col = np.empty(n + 1, np.int32), np.empty(max(2 * k, 1), np.int32)
    val = np.empty(max(2 * k, 1), np.float32) if weighted else None
    t = lib().orc_dense_csr_s(a, n, k, _ptr(row), _ptr(col), _ptr(val))
    return row, col[:2 * k], (val[:2 * k] if weighted else None), t
