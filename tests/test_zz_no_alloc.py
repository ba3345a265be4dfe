"""pg_cp_binary_static (no device allocation) against pg_cp_binary, through
the C ABI: bit-identical clustering coefficients and mean on two calls in a
row (the second with a smaller graph, so stale scratch from the first would
show), rejection of sizes beyond the fixed scratch and of host buffers, and
no change in free device memory across the calls."""
import ctypes as C

import numpy as np
import pytest

from conftest import say
from helpers import random_graph

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _free_bytes(h):
    f, t = C.c_size_t(0), C.c_size_t(0)
    assert h.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def test_static_cp_matches_allocating_entry_point():
    import torch
    import pagani_amd
    lib = pagani_amd.library()
    lib.pg_cp_binary_static.restype = C.c_int
    lib.pg_cp_binary_static.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, C.c_void_p,
                                        C.c_void_p]
    hip = _hip()
    graphs = [random_graph(3000, 0.02, 21), random_graph(700, 0.05, 22)]
    with pagani_amd.Context(0) as ctx:
        # every device buffer exists before the first static call
        dev = []
        for row, col, _ in graphs:
            n = row.size - 1
            dev.append((torch.from_numpy(row).cuda(), torch.from_numpy(col).cuda(),
                        torch.empty(n, dtype=torch.float32, device="cuda"),
                        torch.empty(1, dtype=torch.float64, device="cuda")))
        want = [ctx.cp_binary(row, col) for row, col, _ in graphs]
        torch.cuda.synchronize()
        assert hip.hipDeviceSynchronize() == 0
        free0 = _free_bytes(hip)
        for k, ((row, col, _), (drow, dcol, dcp, dmean)) in enumerate(zip(graphs, dev)):
            n = row.size - 1
            say(f"static cp call {k}: n={n} nnz={col.size}")
            hmean = C.c_double(-1.0)
            st = lib.pg_cp_binary_static(ctx._h, C.c_void_p(drow.data_ptr()), C.c_void_p(dcol.data_ptr()), n,
                                         col.size, C.c_void_p(dcp.data_ptr()), C.cast(C.pointer(hmean), C.c_void_p))
            assert st == 0, lib.pg_last_error(ctx._h)
            cp = dcp.cpu().numpy()
            np.testing.assert_array_equal(cp.view(np.uint32), want[k][0].view(np.uint32))
            assert hmean.value == want[k][1]
            # device destination of the mean
            st = lib.pg_cp_binary_static(ctx._h, C.c_void_p(drow.data_ptr()), C.c_void_p(dcol.data_ptr()), n,
                                         col.size, C.c_void_p(dcp.data_ptr()), C.c_void_p(dmean.data_ptr()))
            assert st == 0
            assert float(dmean.cpu()[0]) == want[k][1]
        assert hip.hipDeviceSynchronize() == 0
        assert _free_bytes(hip) == free0, "device memory was allocated"
        # beyond the fixed scratch: rejected before any launch
        drow, dcol, dcp, _ = dev[0]
        st = lib.pg_cp_binary_static(ctx._h, C.c_void_p(drow.data_ptr()), C.c_void_p(dcol.data_ptr()),
                                     491521, 0, C.c_void_p(dcp.data_ptr()), None)
        assert st == 6  # PG_ENOTSUP
        # host buffers would need staging: rejected
        row, col, _ = graphs[1]
        out = np.empty(row.size - 1, np.float32)
        st = lib.pg_cp_binary_static(ctx._h, row.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p),
                                     row.size - 1, col.size, out.ctypes.data_as(C.c_void_p), None)
        assert st == 1  # PG_EINVAL
        # a malformed CSR is rejected by the check that uses the fixed flag
        bad = dev[1][1].clone()
        bad[5] = 10**6
        drow, _, dcp, _ = dev[1]
        st = lib.pg_cp_binary_static(ctx._h, C.c_void_p(drow.data_ptr()), C.c_void_p(bad.data_ptr()),
                                     row.size - 1, col.size, C.c_void_p(dcp.data_ptr()), None)
        assert st == 4  # PG_EFORMAT
        # and the entry point still works afterwards
        drow, dcol, dcp, _ = dev[1]
        st = lib.pg_cp_binary_static(ctx._h, C.c_void_p(drow.data_ptr()), C.c_void_p(dcol.data_ptr()),
                                     row.size - 1, col.size, C.c_void_p(dcp.data_ptr()), None)
        assert st == 0
        np.testing.assert_array_equal(dcp.cpu().numpy().view(np.uint32), want[1][0].view(np.uint32))
