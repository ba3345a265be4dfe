"""ctypes binding of libpagani.so.

Mirrors the entry points of include/pagani.h one to one; each method names
the reference interface it stands in for (see the header).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PAGANI_LIB") or os.path.join(os.path.dirname(HERE), "lib", "libpagani.so")

PREC_FP32, PREC_BF16X3 = 0, 1
THR_R, THR_SPARSITY = 0, 1
LAYOUT_TIME_MAJOR, LAYOUT_VOXEL_MAJOR = 0, 1
CP_BARRAT, CP_ONNELA = 1, 2
MOD_STRENGTH, MOD_BINARY_K = 0, 1
MOD_STOP_VECTOR, MOD_STOP_RAYLEIGH = 0, 1
# pg_subject.status bits
PG_COHORT_WINDOW, PG_COHORT_SEGMENTS, PG_COHORT_CANDIDATES, PG_COHORT_SELECTION = 1, 2, 4, 8
PG_COHORT_CSR, PG_COHORT_DEPTH, PG_COHORT_OUTSIDE = 16, 32, 64

STATUS = {0: "ok", 1: "invalid argument", 2: "out of device memory", 3: "device error",
          4: "malformed input", 5: "threshold error", 6: "not supported", 7: "bad call order"}


class PgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"pagani status {status} ({STATUS.get(status, '?')}): {msg}")
        self.status = status


class _Opts(C.Structure):
    _fields_ = [("thr_kind", C.c_int32), ("thr", C.c_float), ("weighted", C.c_int32),
                ("prec", C.c_int32), ("layout", C.c_int32)]


class _ModOpts(C.Structure):
    _fields_ = [("kmode", C.c_int32), ("stop", C.c_int32), ("tol", C.c_double), ("max_iter", C.c_int32),
                ("progress", C.c_int32), ("max_depth", C.c_int32)]


class Subject(C.Structure):
    """pg_subject: one subject of pg_cohort_run (device pointers as integers)."""
    _fields_ = [("bold", C.c_void_p), ("n", C.c_int32), ("t", C.c_int32), ("sparsity", C.c_float),
                ("sources", C.c_void_p), ("nsrc", C.c_int64), ("row", C.c_void_p), ("col", C.c_void_p),
                ("deg", C.c_void_p), ("cp", C.c_void_p), ("bc", C.c_void_p), ("status", C.c_uint32),
                ("ms", C.c_float)]


class CohortInfo(C.Structure):
    _fields_ = [("graph_nodes", C.c_int32), ("graphs", C.c_int32), ("reruns", C.c_int32), ("ms_build", C.c_double),
                ("ms_graph", C.c_double), ("ms_reruns", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("value", C.c_int32), ("candidates", C.c_int32), ("ms", C.c_double), ("ms_default", C.c_double)]


class _Info(C.Structure):
    _fields_ = [("n", C.c_int32), ("t", C.c_int32), ("nnz", C.c_int64),
                ("r_thresh", C.c_float), ("s_thresh", C.c_float), ("t_star", C.c_double),
                ("band", C.c_double), ("band_pairs", C.c_int64), ("band_flips", C.c_int64),
                ("gram_passes", C.c_int64), ("ms_zscore", C.c_double), ("ms_gram", C.c_double),
                ("ms_select", C.c_double), ("ms_csr", C.c_double),
                ("ms_gram_hist", C.c_double), ("ms_gram_emit", C.c_double),
                ("band_max_err", C.c_double)]


@dataclass
class CorrInfo:
    n: int
    t: int
    nnz: int
    r_thresh: float
    s_thresh: float
    t_star: float
    band: float
    band_pairs: int
    band_flips: int
    gram_passes: int
    ms_zscore: float
    ms_gram: float
    ms_select: float
    ms_csr: float
    ms_gram_hist: float
    ms_gram_emit: float
    band_max_err: float


_LIB = None
_VP = C.c_void_p

# name -> (restype, argtypes)
_SIGS = {
    "pg_create": (C.c_int, [C.POINTER(_VP), C.c_int]),
    "pg_destroy": (None, [_VP]),
    "pg_last_error": (C.c_char_p, [_VP]),
    "pg_status_string": (C.c_char_p, [C.c_int]),
    "pg_abi_version": (C.c_int, []),
    "pg_synchronize": (C.c_int, [_VP]),
    "pg_set_timing": (C.c_int, [_VP, C.c_int32]),
    "pg_corr_build": (C.c_int, [_VP, _VP, C.c_int32, C.c_int32, C.POINTER(_Opts), C.POINTER(_Info)]),
    "pg_corr_emit": (C.c_int, [_VP, _VP, C.c_int32, C.c_int32, C.POINTER(_Opts), C.POINTER(_Info)]),
    "pg_corr_select": (C.c_int, [_VP, C.c_float, C.POINTER(_Info)]),
    "pg_corr_csr": (C.c_int, [_VP, _VP, _VP, _VP]),
    "pg_sparsity_pairs": (C.c_int64, [C.c_int32, C.c_float]),
    "pg_autotune_corr": (C.c_int, [_VP, _VP, C.c_int32, C.c_int32, C.POINTER(_Opts), C.POINTER(TuneResult)]),
    "pg_autotune_bc": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_int64, _VP, C.c_int64, C.POINTER(TuneResult)]),
    "pg_autotune_fw": (C.c_int, [_VP, _VP, C.c_int32, C.POINTER(TuneResult)]),
    "pg_cohort_run": (C.c_int, [_VP, C.POINTER(Subject), C.c_int32, C.c_int32, C.POINTER(CohortInfo)]),
    "pg_corr_csr_view": (C.c_int, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP)]),
    "pg_corr_z64_view": (C.c_int, [_VP, C.POINTER(_VP)]),
    "pg_degree": (C.c_int, [_VP, _VP, C.c_int32, _VP, _VP]),
    "pg_strength": (C.c_int, [_VP, _VP, _VP, C.c_int32, _VP]),
    "pg_triangles": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_int64, _VP]),
    "pg_cp_binary": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_int64, _VP, _VP]),
    "pg_cp_weighted": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int32, C.c_int64, C.c_int32, C.c_int32, _VP, _VP]),
    "pg_bc": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_int64, _VP, C.c_int64, _VP]),
    "pg_bc_float": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_int64, _VP, C.c_int64, _VP]),
    "pg_bfs": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_int64, _VP, C.c_int64, _VP]),
    "pg_bfs_efficiency": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_int64, _VP, _VP, _VP, _VP]),
    "pg_eigenvector_centrality": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int32, C.c_int64, _VP, _VP, _VP, _VP]),
    "pg_participation": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int32, C.c_int64, _VP, _VP, _VP, _VP]),
    "pg_rewire": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int32, C.c_int64, C.c_uint64, C.c_int64, _VP, _VP, _VP]),
    "pg_cormat_begin": (C.c_int, [_VP, C.c_int32, C.c_int32]),
    "pg_cormat_add": (C.c_int, [_VP, _VP, C.c_int32, C.c_int32]),
    "pg_cormat_finish": (C.c_int, [_VP]),
    "pg_cormat_read": (C.c_int, [_VP, C.c_int64, C.c_int64, _VP]),
    "pg_cormat_threshold": (C.c_int, [_VP, C.c_int32, C.c_float, C.c_int32, _VP]),
    "pg_fw": (C.c_int, [_VP, _VP, C.c_int32, C.c_int32]),
    "pg_fw_init": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int32, C.c_int64, C.c_int32, _VP]),
    "pg_fw_efficiency": (C.c_int, [_VP, _VP, C.c_int32, _VP, _VP, _VP]),
    "pg_modularity": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int32, C.c_int64, C.c_int32, _VP, _VP, _VP, _VP, _VP]),
    "pg_modularity_ex": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int32, C.c_int64, C.POINTER(_ModOpts), _VP, _VP, _VP,
                                   _VP, _VP]),
}


def library():
    """Load libpagani.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run python pagani-toolkit_amd/build.py")
        # PyTorch ships its own HIP runtime with the same SONAME
        # (libamdhip64.so.7) but asks for it by the file name libamdhip64.so.
        # Loaded first, it also serves libpagani; loaded after libpagani has
        # pulled in ROCm's copy, it becomes a second runtime that finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB


def _ptr(x):
    """Host numpy array or torch tensor -> void*."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return x.ctypes.data_as(_VP)
    if hasattr(x, "data_ptr"):
        if not x.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return _VP(x.data_ptr())
    raise TypeError(f"unsupported array type {type(x)!r}")


def _len(x):
    return int(x.numel()) if hasattr(x, "numel") else int(x.size)


class Context:
    """One pg_ctx: a HIP stream plus a reusable device arena."""

    def __init__(self, device: int = 0):
        self._lib = library()
        h = _VP()
        st = self._lib.pg_create(C.byref(h), device)
        if st != 0:
            raise PgError(st, "pg_create failed (no usable HIP device?)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != 0:
            msg = self._lib.pg_last_error(self._h)
            raise PgError(st, msg.decode() if msg else "")

    def synchronize(self):
        self._check(self._lib.pg_synchronize(self._h))

    def set_timing(self, level: int):
        """Stage timing of corr_build: 2 every stage (default), 1 the
        emitting Gram kernel only, 0 none (untimed stages report 0)."""
        self._check(self._lib.pg_set_timing(self._h, int(level)))

    # ------------------------------------------------------- construction
    def corr_build(self, bold, n: int, t: int, thr: float, *, sparsity: bool = False,
                   weighted: bool = False, prec: int = PREC_BF16X3,
                   layout: int = LAYOUT_TIME_MAJOR) -> CorrInfo:
        """CorMat_gpu + post_block + FormFullGraph(_s) + FormCSRGraph."""
        o = _Opts(THR_SPARSITY if sparsity else THR_R, thr, int(weighted), prec, layout)
        info = _Info()
        self._check(self._lib.pg_corr_build(self._h, _ptr(bold), n, t, C.byref(o), C.byref(info)))
        return CorrInfo(**{f: getattr(info, f) for f, _ in _Info._fields_})

    def corr_emit(self, bold, n: int, t: int, thr: float, *, sparsity: bool = False,
                  weighted: bool = False, prec: int = PREC_BF16X3,
                  layout: int = LAYOUT_TIME_MAJOR) -> CorrInfo:
        """One Gram pass keeping every pair the loosest threshold `thr` can keep."""
        o = _Opts(THR_SPARSITY if sparsity else THR_R, thr, int(weighted), prec, layout)
        info = _Info()
        self._check(self._lib.pg_corr_emit(self._h, _ptr(bold), n, t, C.byref(o), C.byref(info)))
        return CorrInfo(**{f: getattr(info, f) for f, _ in _Info._fields_})

    def corr_select(self, thr: float) -> CorrInfo:
        """CSR of one threshold (as strict as the emitted one or stricter) from the kept pairs."""
        info = _Info()
        self._check(self._lib.pg_corr_select(self._h, thr, C.byref(info)))
        return CorrInfo(**{f: getattr(info, f) for f, _ in _Info._fields_})

    def corr_csr(self, n: int, nnz: int, weighted: bool = False):
        row = np.empty(n + 1, np.int32)
        col = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1), np.float32) if weighted else None
        self._check(self._lib.pg_corr_csr(self._h, _ptr(row), _ptr(col), _ptr(val)))
        return row, col[:nnz], (val[:nnz] if weighted else None)

    def corr_csr_view(self):
        r, c, v = _VP(), _VP(), _VP()
        self._check(self._lib.pg_corr_csr_view(self._h, C.byref(r), C.byref(c), C.byref(v)))
        return r.value, c.value, v.value

    def corr_z64_view(self):
        z = _VP()
        self._check(self._lib.pg_corr_z64_view(self._h, C.byref(z)))
        return z.value

    # ---------------------------------------------------------- autotuner
    def autotune_corr(self, bold, n: int, t: int, thr: float, *, sparsity: bool = False, weighted: bool = False,
                      prec: int = PREC_BF16X3, layout: int = LAYOUT_TIME_MAJOR) -> TuneResult:
        """Gram supertile edge for (arch, n, t): cached, or searched on this input."""
        o = _Opts(THR_SPARSITY if sparsity else THR_R, thr, int(weighted), prec, layout)
        r = TuneResult()
        self._check(self._lib.pg_autotune_corr(self._h, _ptr(bold), n, t, C.byref(o), C.byref(r)))
        return r

    def autotune_bc(self, row, col, sources) -> TuneResult:
        """64-source BFS batches per launch for (arch, n, nnz)."""
        r = TuneResult()
        self._check(self._lib.pg_autotune_bc(self._h, _ptr(row), _ptr(col), _len(row) - 1, _len(col), _ptr(sources),
                                             _len(sources), C.byref(r)))
        return r

    def autotune_fw(self, d) -> TuneResult:
        """Floyd-Warshall tile for (arch, n); d (n x n) is not changed."""
        r = TuneResult()
        self._check(self._lib.pg_autotune_fw(self._h, _ptr(d), int(d.shape[0]), C.byref(r)))
        return r

    # ------------------------------------------------------------- cohort
    def sparsity_pairs(self, n: int, sparsity: float) -> int:
        return int(self._lib.pg_sparsity_pairs(n, sparsity))

    def cohort_run(self, subjects, launches: int = 1) -> CohortInfo:
        """pg_cohort_run over a list of Subject structures (updated in place:
        status and ms)."""
        arr = (Subject * len(subjects))(*subjects)
        info = CohortInfo()
        self._check(self._lib.pg_cohort_run(self._h, arr, len(subjects), launches, C.byref(info)))
        for k in range(len(subjects)):
            subjects[k].status = arr[k].status
            subjects[k].ms = arr[k].ms
        return info

    # ------------------------------------------------- dense matrix, averaging
    def cormat_begin(self, n: int, fisher: bool = False):
        self._check(self._lib.pg_cormat_begin(self._h, n, int(fisher)))

    def cormat_add(self, bold, t: int, layout: int = LAYOUT_TIME_MAJOR):
        self._check(self._lib.pg_cormat_add(self._h, _ptr(bold), t, layout))

    def cormat_finish(self):
        self._check(self._lib.pg_cormat_finish(self._h))

    def cormat_read(self, n: int, offset: int = 0, count=None, out=None):
        m1 = n * (n - 1) // 2
        count = m1 - offset if count is None else count
        out = np.empty(count, np.float32) if out is None else out
        self._check(self._lib.pg_cormat_read(self._h, offset, count, _ptr(out)))
        return out

    def cormat_threshold(self, thr: float, *, sparsity: bool = False, weighted: bool = False) -> CorrInfo:
        info = _Info()
        self._check(self._lib.pg_cormat_threshold(self._h, THR_SPARSITY if sparsity else THR_R, thr,
                                                   int(weighted), C.byref(info)))
        return CorrInfo(**{f: getattr(info, f) for f, _ in _Info._fields_})

    # ------------------------------------------------------------ metrics
    def degree(self, row, deg=None, deg_ref=None):
        n = _len(row) - 1
        if deg is None and deg_ref is None:
            deg, deg_ref = np.empty(n, np.int32), np.empty(n, np.float32)
        self._check(self._lib.pg_degree(self._h, _ptr(row), n, _ptr(deg), _ptr(deg_ref)))
        return deg, deg_ref

    def strength(self, row, val, out=None):
        n = _len(row) - 1
        out = np.empty(n, np.float32) if out is None else out
        self._check(self._lib.pg_strength(self._h, _ptr(row), _ptr(val), n, _ptr(out)))
        return out

    def triangles(self, row, col, out=None):
        n = _len(row) - 1
        out = np.empty(n, np.int64) if out is None else out
        self._check(self._lib.pg_triangles(self._h, _ptr(row), _ptr(col), n, _len(col), _ptr(out)))
        return out

    def cp_binary(self, row, col, out=None):
        n = _len(row) - 1
        out = np.empty(n, np.float32) if out is None else out
        mean = C.c_double(0.0)
        self._check(self._lib.pg_cp_binary(self._h, _ptr(row), _ptr(col), n, _len(col), _ptr(out),
                                           C.cast(C.pointer(mean), _VP)))
        return out, mean.value

    def cp_weighted(self, row, col, val, kind: int = CP_ONNELA, wmax_mode: int = 0, out=None):
        n = _len(row) - 1
        out = np.empty(n, np.float64) if out is None else out
        mean = C.c_double(0.0)
        self._check(self._lib.pg_cp_weighted(self._h, _ptr(row), _ptr(col), _ptr(val), n, _len(col),
                                             kind, wmax_mode, _ptr(out), C.cast(C.pointer(mean), _VP)))
        return out, mean.value

    def bc(self, row, col, sources=None, out=None):
        n = _len(row) - 1
        out = np.empty(n, np.float64) if out is None else out
        ns = 0 if sources is None else _len(sources)
        self._check(self._lib.pg_bc(self._h, _ptr(row), _ptr(col), n, _len(col), _ptr(sources), ns, _ptr(out)))
        return out

    def bc_float(self, row, col, sources=None, out=None):
        """pg_bc rounded to float32, the output type of CUBC."""
        n = _len(row) - 1
        out = np.empty(n, np.float32) if out is None else out
        ns = 0 if sources is None else _len(sources)
        self._check(self._lib.pg_bc_float(self._h, _ptr(row), _ptr(col), n, _len(col), _ptr(sources), ns,
                                          _ptr(out)))
        return out

    def bfs(self, row, col, sources, out=None):
        n = _len(row) - 1
        ns = _len(sources)
        out = np.empty((ns, n), np.int32) if out is None else out
        self._check(self._lib.pg_bfs(self._h, _ptr(row), _ptr(col), n, _len(col), _ptr(sources), ns, _ptr(out)))
        return out

    def bfs_efficiency(self, row, col, eff=None):
        """(eff[n] float32, Lp, isolated, unconnected) of a binary network."""
        n = _len(row) - 1
        eff = np.empty(n, np.float32) if eff is None else eff
        lp, iso, unc = C.c_double(0), C.c_int32(0), C.c_int32(0)
        self._check(self._lib.pg_bfs_efficiency(self._h, _ptr(row), _ptr(col), n, _len(col), _ptr(eff),
                                                 C.cast(C.pointer(lp), _VP), C.cast(C.pointer(iso), _VP),
                                                 C.cast(C.pointer(unc), _VP)))
        return eff, lp.value, iso.value, unc.value

    def eigenvector_centrality(self, row, col, val=None):
        """(vhat, avec, eigenvalue, iterations) of the reference's power iteration."""
        n = _len(row) - 1
        vh, av = np.empty(n, np.float64), np.empty(n, np.float64)
        lam, it = C.c_double(0), C.c_int32(0)
        self._check(self._lib.pg_eigenvector_centrality(self._h, _ptr(row), _ptr(col), _ptr(val), n, _len(col),
                                                         _ptr(vh), _ptr(av), C.cast(C.pointer(lam), _VP),
                                                         C.cast(C.pointer(it), _VP)))
        return vh, av, lam.value, it.value

    def participation(self, row, col, labels, val=None):
        """(pc, pc_normalized) float32 and the module count K."""
        n = _len(row) - 1
        p, q = np.empty(n, np.float32), np.empty(n, np.float32)
        k = C.c_int32(0)
        self._check(self._lib.pg_participation(self._h, _ptr(row), _ptr(col), _ptr(val), n, _len(col),
                                                _ptr(labels), _ptr(p), _ptr(q), C.cast(C.pointer(k), _VP)))
        return p, q, k.value

    def rewire(self, row, col, val=None, seed: int = 0, attempts: int = 0, out_col=None, out_val=None):
        """Degree-preserving random rewiring: (col, val or None, swaps made).

        The row offsets are unchanged; out_col / out_val may be device tensors."""
        nnz = _len(col)
        oc = np.empty(nnz, np.int32) if out_col is None else out_col
        ov = None if val is None else (np.empty(nnz, np.float32) if out_val is None else out_val)
        sw = C.c_int64(0)
        self._check(self._lib.pg_rewire(self._h, _ptr(row), _ptr(col), _ptr(val), _len(row) - 1, nnz,
                                        seed & (2**64 - 1), attempts, _ptr(oc), _ptr(ov),
                                        C.cast(C.pointer(sw), _VP)))
        return oc, ov, sw.value

    def fw(self, d, block: int = 0):
        """Closure in place; d is an n x n float32 array/tensor."""
        n = int(d.shape[0])
        self._check(self._lib.pg_fw(self._h, _ptr(d), n, block))
        return d

    def fw_init(self, row, col, val, kind: int, out=None):
        n = _len(row) - 1
        out = np.empty((n, n), np.float32) if out is None else out
        self._check(self._lib.pg_fw_init(self._h, _ptr(row), _ptr(col), _ptr(val), n, _len(col), kind, _ptr(out)))
        return out

    def fw_efficiency(self, d, eff=None):
        """(eff[n] float32, Lp, isolated count) from a closed distance matrix."""
        n = int(d.shape[0])
        eff = np.empty(n, np.float32) if eff is None else eff
        lp, iso = C.c_double(0), C.c_int32(0)
        self._check(self._lib.pg_fw_efficiency(self._h, _ptr(d), n, _ptr(eff), C.cast(C.pointer(lp), _VP),
                                                C.cast(C.pointer(iso), _VP)))
        return eff, lp.value, iso.value

    def modularity(self, row, col, val=None, kmode: int = MOD_STRENGTH, stop: int = MOD_STOP_VECTOR,
                   tol: float = 0.0, max_iter: int = 0, progress: bool = False, max_depth: int = 0):
        """(labels, Q, lambda0, round-0 vector, rounds): Partition_GPU with the
        reference's stopping rule (MOD_STOP_VECTOR) or the Rayleigh-quotient rule."""
        n = _len(row) - 1
        labels = np.empty(n, np.int32)
        eig = np.empty(n, np.float64)
        q, lam, nr = C.c_double(0), C.c_double(0), C.c_int32(0)
        o = _ModOpts(kmode, stop, tol, max_iter, int(progress), max_depth)
        self._check(self._lib.pg_modularity_ex(self._h, _ptr(row), _ptr(col), _ptr(val), n, _len(col), C.byref(o),
                                               _ptr(labels), C.cast(C.pointer(q), _VP),
                                               C.cast(C.pointer(lam), _VP), _ptr(eig),
                                               C.cast(C.pointer(nr), _VP)))
        return labels, q.value, lam.value, eig, nr.value
