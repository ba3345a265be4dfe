"""pg_cohort_run: uneven subjects in one captured graph give exactly what the
ordinary entry points give subject by subject (CSR, degree and clustering
bit-exact, BC to FP64 summation order), also on graph replays; subjects the
graph cannot finish are rebuilt outside it and flagged."""
import numpy as np
import pytest

from conftest import say
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(6000, 150), (9000, 300), (7500, 200), (12000, 240), (6500, 170), (1500, 120)]


@pytest.fixture(scope="module")
def cohort_data():
    from pagani_amd import synth
    return [(n, t, synth.bold(4000 + s, n, t, 90), synth.sample_sources(n, 256, 4242))
            for s, (n, t) in enumerate(SHAPES)]


def _subjects(ctx, data, sparsity=1.0):
    import torch
    import pagani_amd
    subs, keep = [], []
    for n, t, x, src in data:
        k = ctx.sparsity_pairs(n, sparsity)
        b = torch.from_numpy(x).cuda()
        row = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        col = torch.empty(max(2 * k, 1), dtype=torch.int32, device="cuda")
        deg = torch.empty(n, dtype=torch.int32, device="cuda")
        cp = torch.empty(n, dtype=torch.float32, device="cuda")
        bc = torch.empty(n, dtype=torch.float64, device="cuda")
        s = pagani_amd.Subject(b.data_ptr(), n, t, sparsity, src.ctypes.data, src.size, row.data_ptr(),
                               col.data_ptr(), deg.data_ptr(), cp.data_ptr(), bc.data_ptr(), 0, 0.0)
        subs.append(s)
        keep.append((b, row, col, deg, cp, bc, k))
    return subs, keep


def _check_against_entry_points(ctx, data, keep, sparsity=1.0):
    for (n, t, x, src), (_, row, col, deg, cp, bc, k) in zip(data, keep):
        info = ctx.corr_build(x, n, t, sparsity, sparsity=True)
        erow, ecol, _ = ctx.corr_csr(n, info.nnz)
        assert info.nnz == 2 * k
        np.testing.assert_array_equal(row.cpu().numpy(), erow)
        np.testing.assert_array_equal(col.cpu().numpy()[:2 * k], ecol)
        edeg, _ = ctx.degree(erow)
        np.testing.assert_array_equal(deg.cpu().numpy(), edeg)
        ecp, _ = ctx.cp_binary(erow, ecol)
        np.testing.assert_array_equal(cp.cpu().numpy().view(np.uint32), ecp.view(np.uint32))
        ebc = ctx.bc(erow, ecol, src)
        np.testing.assert_allclose(bc.cpu().numpy(), ebc, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("launches", [1, 3])
def test_cohort_equals_entry_points(ctx, cohort_data, launches):
    import pagani_amd
    subs, keep = _subjects(ctx, cohort_data)
    say(f"cohort {len(subs)} subjects, launches={launches}")
    info = ctx.cohort_run(subs, launches)
    print(f"cohort: {info.graphs} graph(s), {info.graph_nodes} nodes, build {info.ms_build:.1f} ms, "
          f"graph {info.ms_graph:.2f} ms, "
          f"{info.reruns} reruns; per subject ms {[round(s.ms, 3) for s in subs]}")
    assert [s.status for s in subs[:-1]] == [0] * (len(subs) - 1)
    assert subs[-1].status == pagani_amd.lib.PG_COHORT_OUTSIDE  # too few supertiles for the single pass
    assert info.reruns == 1
    assert info.graph_nodes > 0 and info.ms_graph > 0 and info.graphs == 1
    _check_against_entry_points(ctx, cohort_data, keep)
    # and the first subject against the FP64 oracle
    n, t, x, _ = cohort_data[0]
    orow, ocol, _, _ = O.corr_csr_s(O.zscore(x)[0], 1.0)
    np.testing.assert_array_equal(keep[0][1].cpu().numpy(), orow)
    np.testing.assert_array_equal(keep[0][2].cpu().numpy(), ocol)


def test_cohort_split_into_several_graphs(ctx, cohort_data, monkeypatch):
    # a node budget below one subject's launches: one graph per subject,
    # launched back to back, same results
    monkeypatch.setenv("PAGANI_COHORT_GRAPH_NODES", "10")
    data = cohort_data[:4]
    subs, keep = _subjects(ctx, data)
    say("cohort, one graph per subject")
    info = ctx.cohort_run(subs, 2)
    assert info.graphs == 4, info.graphs
    assert [s.status for s in subs] == [0] * 4
    _check_against_entry_points(ctx, data, keep)


@pytest.mark.parametrize("knob,value,bit", [("PAGANI_COHORT_SEG_SCALE", "0.001", 2),
                                            ("PAGANI_COHORT_LEVELS", "2", 32)])
def test_cohort_flags_and_rebuilds(ctx, cohort_data, monkeypatch, knob, value, bit):
    # segments far too small, or too few captured BFS levels: the device
    # checks flag the subjects and the ordinary entry points rebuild them
    monkeypatch.setenv(knob, value)
    data = cohort_data[:3]
    subs, keep = _subjects(ctx, data)
    say(f"cohort with {knob}={value}")
    info = ctx.cohort_run(subs, 1)
    assert all(s.status & bit for s in subs), [s.status for s in subs]
    assert info.reruns == len(subs)
    _check_against_entry_points(ctx, data, keep)
