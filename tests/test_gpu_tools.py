"""GPU: the command-line tools end to end on synthetic NIfTI volumes.

Files are compared with the FP64 oracle: edge sets and degrees bit-exact,
BC / clustering / efficiency within the tolerances of the library tests.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import say
from oracle import oracle as O
from pagani_amd import formats, synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pagani-toolkit_amd", "bin")


def run(tool, *args, cwd=None, env=None):
    say(f"run {tool} {' '.join(map(str, args))}")
    r = subprocess.run([os.path.join(BIN, tool), *map(str, args)], capture_output=True, text=True,
                       timeout=600, cwd=cwd, env=None if env is None else {**os.environ, **env})
    assert r.returncode == 0, (tool, r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


@pytest.fixture(scope="module")
def volume(tmp_path_factory):
    d = tmp_path_factory.mktemp("bold")
    v, t = 1400, 120
    mask = np.ones(v, np.uint8)
    mask[::17] = 0
    formats.write_nifti(str(d / "mask.nii"), mask, formats.DT_UINT8)
    subs = {}
    for name, seed, dt in [("sub01", 21, formats.DT_FLOAT32), ("sub02", 22, formats.DT_FLOAT64)]:
        x = synth.bold(seed, v, t, 8)
        formats.write_nifti(str(d / f"{name}.nii"), x if dt == formats.DT_FLOAT32 else x.astype(np.float64), dt)
        subs[name] = np.ascontiguousarray(x[:, mask >= 1])
    return d, subs, int((mask >= 1).sum())


def test_cucormat_rmode_files(volume):
    d, subs, n = volume
    run("CUCorMat", d, 0.5, "n", "n", "r", 0.35, 0.3)
    for name, x in subs.items():
        z64, _ = O.zscore(x)
        for thr in (0.3, 0.35):
            orow, ocol, _ = O.corr_csr_r(z64, thr)
            m1 = n * (n - 1) // 2
            s = np.float32(100.0 * orow[-1] / m1 / 2)
            path = d / formats.csr_name(f"{name}", n, float(s), float(np.float32(thr)))
            row, col, val = formats.read_csr(str(path))
            assert val is None
            np.testing.assert_array_equal(row, orow)
            np.testing.assert_array_equal(col, ocol)


def test_cucormat_documented_form_sparsity_weighted(volume):
    d, subs, n = volume
    run("CUCorMat", "--weighted", d, d / "mask.nii", 0.5, "n", "n", "s", 2.0)
    out = d / "weighted"
    for name, x in subs.items():
        z64, _ = O.zscore(x)
        orow, ocol, oval, tstar = O.corr_csr_s(z64, 2.0, True)
        path = out / formats.csr_name(name, n, 2.0, float(np.float32(tstar)), True)
        row, col, val = formats.read_csr(str(path))
        np.testing.assert_array_equal(row, orow)
        np.testing.assert_array_equal(col, ocol)
        np.testing.assert_array_equal(val.view(np.uint32), oval.view(np.uint32))


def _csr_files(d, weighted=False):
    return sorted(f for f in os.listdir(d) if f.endswith(".csr"))


def _ensure_networks(d):
    """The per-subject networks the metric tools read (the first two tests
    write them; a selection that skips those tests gets them here)."""
    if not _csr_files(d):
        run("CUCorMat", d, 0.5, "n", "n", "r", 0.35, 0.3)
    if not (d / "weighted").exists() or not _csr_files(d / "weighted"):
        run("CUCorMat", "--weighted", d, d / "mask.nii", 0.5, "n", "n", "s", 2.0)


def test_metric_tools_on_binary_networks(volume):
    d, subs, n = volume
    _ensure_networks(d)
    run("Degree", d)
    run("CUBC", d)
    run("Cp", d, 0, "gn")
    for f in _csr_files(d):
        base = str(d / f)[:-4]
        row, col, _ = formats.read_csr(str(d / f))
        _, dref = O.degree(row)
        np.testing.assert_array_equal(formats.read_nm(base + "_deg.nm"), dref)
        np.testing.assert_allclose(formats.read_nm(base + "_bc.nm"), O.bc(row, col).astype(np.float32), rtol=1e-6)
        cp, mean = O.cp_binary(row, col)
        np.testing.assert_array_equal(formats.read_nm(base + "_cp.nm"), cp)
        assert abs(float(open(base + "_cp.txt").read()) - mean) < 1e-5
    iso = open(d / "isolated_v_mark.txt").read().split()
    assert len(iso) == len(_csr_files(d))


def test_bfs_lp_tool_on_binary_networks(volume):
    d, subs, n = volume
    _ensure_networks(d)
    run("CUBFS_Lp", d, 0, "b")
    for f in _csr_files(d):
        base = str(d / f)[:-4]
        row, col, _ = formats.read_csr(str(d / f))
        eff, lp, _, _ = O.bfs_efficiency(row, col)
        np.testing.assert_array_equal(formats.read_nm(base + "_eff.nm").view(np.uint32), eff.view(np.uint32))
        assert abs(float(open(base + "_Lp.txt").read()) - lp) <= 5e-6 * lp  # six significant digits


def test_lp_and_modularity_tools_on_weighted_networks(volume):
    d, subs, n = volume
    _ensure_networks(d)
    w = d / "weighted"
    run("CUBFW_Lp", w, 0, "gn")
    run("CUmodularity", w, 0)
    run("Degree", "--weighted", w)
    run("Cp", "--weighted", w, 0, 2, "n")
    for f in _csr_files(w):
        base = str(w / f)[:-4]
        row, col, val = formats.read_csr(str(w / f))
        # efficiency from the oracle's blocked closure with the tool's tile edge (64)
        dist = O.fw_blocked(O.fw_init(row, col, val, 0), 64)
        inv = np.where((dist >= 0.1) & ~np.eye(n, dtype=bool), 1.0 / dist.astype(np.float64), 0.0)
        eff = (inv.sum(1) / (n - 1)).astype(np.float32)
        eff[eff < np.float32(1.0 / n)] = 0
        np.testing.assert_allclose(formats.read_nm(base + "_eff.nm"), eff, rtol=1e-5, atol=1e-7)
        lp = 1.0 / (inv.sum() / n / (n - 1))
        assert abs(float(open(base + "_Lp.txt").read()) - lp) <= 1e-5 * lp
        np.testing.assert_array_equal(formats.read_nm(base + "_deg.nm"), O.strength(row, val))
        cp, _ = O.cp_weighted(row, col, val, 2, 0)
        np.testing.assert_allclose(formats.read_nm(base + "_cp.nm"), cp.astype(np.float32), rtol=1e-6)
        lab, q, *_ = O.modularity(row, col, val, 1)
        got = formats.read_nm(base + "_modu.nm").astype(np.int32)
        assert (got == lab).mean() > 0.99
        txt = open(base + "_modu.txt").read()
        qfile = float(txt.split("Q=")[1].split()[0])
        assert abs(qfile - q) <= 1e-5


def test_ec_and_pc_tools(volume):
    d, subs, n = volume
    _ensure_networks(d)
    w = d / "weighted"
    run("CUmodularity", w, 0)      # module labels for PC_CPU
    run("PC_CPU", w)
    run("CUEC", "--weighted", w)
    run("CUEC", d)
    for f in _csr_files(w):
        base = str(w / f)[:-4]
        row, col, val = formats.read_csr(str(w / f))
        lab = formats.read_nm(base + "_modu.nm").astype(np.int32)
        p, q, _ = O.pc(row, col, lab, val)
        np.testing.assert_array_equal(formats.read_nm(base + "_pc.nm").view(np.uint32), p.view(np.uint32))
        np.testing.assert_array_equal(formats.read_nm(base + "_pc.nm_Normalized.nm").view(np.uint32),
                                      q.view(np.uint32))
        vh, av, lam, it = O.ec(row, col, val)
        np.testing.assert_allclose(formats.read_nm(base + "_ec.nm"), av.astype(np.float32), rtol=1e-6)
        assert abs(float(open(base + "_evalue.txt").read()) - lam) <= 1e-5 * lam
    for f in _csr_files(d):
        base = str(d / f)[:-4]
        row, col, _ = formats.read_csr(str(d / f))
        vh, av, lam, it = O.ec(row, col)
        np.testing.assert_allclose(formats.read_nm(base + "_ec.nm"), vh.astype(np.float32), rtol=1e-6,
                                   atol=1e-9)
        assert abs(float(open(base + "_evalue.txt").read()) - lam) <= 1e-5 * lam


def _read_cormat(path, n):
    raw = open(path, "rb").read()
    m1 = n * (n - 1) // 2
    assert np.frombuffer(raw[:4], np.int32)[0] == np.int32(np.int64(m1) & 0xffffffff)
    a = np.frombuffer(raw[4:], np.float32)
    assert a.size == m1
    return a


def test_cucormat_average_and_cormat(volume, tmp_path):
    d, subs, n = volume
    src = tmp_path / "grp"
    src.mkdir()
    for f in ["mask.nii", "sub01.nii", "sub02.nii"]:
        (src / f).write_bytes((d / f).read_bytes())
    run("CUCorMat", src, 0.5, "bf", "y", "r", 0.3)
    xs = [subs["sub01"], subs["sub02"]]
    # per-subject dense matrices: (float) of the FP64 r, bit-exact
    for name, x in subs.items():
        a = _read_cormat(src / f"{name}_{n}.cormat", n)
        np.testing.assert_array_equal(a.view(np.uint32), O.cormat_average([x]).view(np.uint32))
    # the Fisher average: within one FP32 ulp of the oracle
    base = src / f"grp_2_{n}"
    avg = _read_cormat(str(base) + ".cormat", n)
    want = O.cormat_average(xs, True)
    assert np.max(np.abs(avg.astype(np.float64) - want)) <= 2 * np.spacing(np.float32(1))
    # its network: exactly the reference's threshold rule on the written matrix
    orow, ocol, _ = O.dense_csr_r(avg, n, 0.3)
    m1 = n * (n - 1) // 2
    name = f"grp_2_{n}_spa{np.float32(100.0 * orow[-1] / m1 / 2):.3f}%_cor{np.float32(0.3):.3f}.csr"
    row, col, val = formats.read_csr(str(src / name))
    assert val is None
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    # weighted sparsity form of the plain average: weights are the matrix values
    run("CUCorMat", "--weighted", src, "0.5", "yn", "n", "s", 1.5)
    avg = O.cormat_average(xs)
    orow, ocol, oval, t = O.dense_csr_s(avg, n, 1.5, True)
    name = f"grp_2_{n}_spa{1.5:.3f}%_cor{t:.3f}.csr"
    row, col, val = formats.read_csr(str(src / name))
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(val.view(np.uint32), oval.view(np.uint32))


def _lines(path):
    return [float(x) for x in open(path).read().split()]


def test_random_network_outputs(volume, tmp_path, ctx):
    # Random networks: network l is the library's rewiring with seed
    # PAGANI_SEED + l, so every line the tools write can be recomputed here.
    d, subs, n = volume
    _ensure_networks(d)
    env = {"PAGANI_SEED": "11"}
    b = tmp_path / "bin"
    b.mkdir()
    f = _csr_files(d)[0]
    (b / f).write_bytes((d / f).read_bytes())
    base = str(b / f)[:-4]
    row, col, _ = formats.read_csr(str(b / f))
    rnd = [ctx.rewire(row, col, seed=11 + l)[0] for l in range(3)]
    for oc in rnd:
        assert O.cp_binary(row, oc)[1] < O.cp_binary(row, col)[1]
    run("Cp", b, 3, "gk", env=env)
    cps = _lines(base + "_cp.txt")
    want = [O.cp_binary(row, col)[1]] + [O.cp_binary(row, oc)[1] for oc in rnd]
    np.testing.assert_allclose(cps, want, rtol=5e-6)
    assert abs(_lines(base + "_gamma.txt")[0] / (want[0] / np.mean(want[1:])) - 1) < 5e-6
    run("CUBFS_Lp", b, 2, "bl", env=env)
    lps = _lines(base + "_Lp.txt")
    want = [O.bfs_efficiency(row, col)[1]] + [O.bfs_efficiency(row, oc)[1] for oc in rnd[:2]]
    np.testing.assert_allclose(lps, want, rtol=5e-6)
    assert abs(_lines(base + "_lambda.txt")[0] / (want[0] / np.mean(want[1:])) - 1) < 5e-6
    run("SmallWorldProperty", b)  # sigma = gamma / lambda from the two files above
    sigma = _lines(base + "_gamma.txt")[0] / _lines(base + "_lambda.txt")[0]
    assert abs(_lines(base + "_smallWorld.txt")[0] / sigma - 1) < 5e-6
    run("Maslov", b / f, env=env)
    r2, c2, v2 = formats.read_csr(base + "_Maslov.csr")
    assert v2 is None
    np.testing.assert_array_equal(r2, row)
    np.testing.assert_array_equal(c2, rnd[0])

    w = tmp_path / "w"
    w.mkdir()
    f = _csr_files(d / "weighted")[0]
    (w / f).write_bytes((d / "weighted" / f).read_bytes())
    base = str(w / f)[:-4]
    row, col, val = formats.read_csr(str(w / f))
    oc, ov, _ = ctx.rewire(row, col, val, seed=11)
    run("CUBFW_Lp", w, 1, "gl", env=env)
    lps = _lines(base + "_Lp.txt")
    want = []
    for c, v in ((col, val), (oc, ov)):
        dist = O.fw_blocked(O.fw_init(row, c, v, 0), 64)
        inv = np.where((dist >= 0.1) & ~np.eye(n, dtype=bool), 1.0 / dist.astype(np.float64), 0.0)
        want.append(1.0 / (inv.sum() / n / (n - 1)))
    np.testing.assert_allclose(lps, want, rtol=1e-5)
    assert abs(_lines(base + "_lambda.txt")[0] / (want[0] / want[1]) - 1) < 2e-5
    run("Cp", "--weighted", w, 1, 2, "gk", env=env)
    want = [O.cp_weighted(row, col, val, 2, 0)[1], O.cp_weighted(row, oc, ov, 2, 0)[1]]
    np.testing.assert_allclose(_lines(base + "_cp.txt"), want, rtol=1e-5)
    run("CUmodularity", w, 1, env=env)
    txt = open(base + "_modu.txt").read()
    q = float(txt.split("Random network 1:")[1].split("Q=")[1].split()[0])
    assert 0.0 < q < float(txt.split("Q=")[1].split()[0])  # a rewired network is less modular
    run("Maslov", w / f, env=env)
    r2, c2, v2 = formats.read_csr(base + "_Maslov.csr")
    np.testing.assert_array_equal(r2, row)
    np.testing.assert_array_equal(c2, oc)
    np.testing.assert_array_equal(v2.view(np.uint32), ov.view(np.uint32))
