"""File formats (Python mirror) and the tools' argument handling, no GPU."""
import os
import struct
import subprocess

import numpy as np
import pytest

from pagani_amd import formats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pagani-toolkit_amd", "bin")


def test_nifti_header_words(tmp_path):
    p = tmp_path / "a.nii"
    data = np.arange(24, dtype=np.float32).reshape(3, 8)
    formats.write_nifti(str(p), data, formats.DT_FLOAT32)
    raw = p.read_bytes()
    h = np.frombuffer(raw[:352], np.int16)
    assert (h[21], h[22], h[23], h[24], h[35], h[36]) == (8, 1, 1, 3, 16, 32)
    np.testing.assert_array_equal(np.frombuffer(raw[352:], np.float32), data.ravel())


def test_csr_and_nm_round_trip(tmp_path):
    row = np.array([0, 2, 3, 5], np.int32)
    col = np.array([1, 2, 0, 0, 1], np.int32)
    val = np.array([0.5, 0.25, 0.5, 0.25, 0.75], np.float32)
    p = tmp_path / "g.csr"
    formats.write_csr(str(p), row, col, val)
    raw = p.read_bytes()
    assert struct.unpack_from("<i", raw, 0)[0] == 4
    assert len(raw) == 4 + 16 + 4 + 20 + 4 + 20
    r, c, v = formats.read_csr(str(p))
    np.testing.assert_array_equal(r, row)
    np.testing.assert_array_equal(c, col)
    np.testing.assert_array_equal(v, val)
    formats.write_csr(str(p), row, col)
    assert formats.read_csr(str(p))[2] is None
    formats.write_nm(str(tmp_path / "x.nm"), [1.5, 2.5])
    np.testing.assert_array_equal(formats.read_nm(str(tmp_path / "x.nm")), [1.5, 2.5])


def test_csr_name_matches_reference_printf():
    # sprintf_s(sparsity, "_spa%.3f%%_cor%.3f", s_thresh, r_thresh) with floats
    s, r = np.float32(1.0), np.float32(0.2139584882412609)
    assert formats.csr_name("sub01", 58000, float(s), float(r)) == "sub01_58000_spa1.000%_cor0.214.csr"
    assert formats.csr_name("sub01", 5, 2.5, 0.4, True) == "sub01_5_spa2.500%_cor0.400_weighted.csr"


@pytest.mark.parametrize("tool,args", [
    ("CUCorMat", []), ("CUBC", []), ("Degree", []), ("Cp", ["x"]), ("CUBFW_Lp", ["x", "0"]),
    ("CUmodularity", ["x"]), ("CUBFS_Lp", ["x", "0"]), ("CUBFS_Lp", ["/nonexistent", "-2", "b"]),
    ("ConvertNII", ["x"]), ("ConvertNII", ["x", "y", "0.5", "z"]), ("CUEC", []), ("PC_CPU", []),
    ("PC_CPU", ["/nonexistent"]),
    ("CUCorMat", ["/nonexistent", "0.5", "y", "n", "r", "0.3"]),   # averaging needs f or n
    ("CUCorMat", ["/nonexistent", "0.5", "q", "n", "r", "0.3"]),   # bad averaging mode
    ("CUCorMat", ["/nonexistent", "0.5", "n", "q", "r", "0.3"]),   # bad cormat flag
    ("CUCorMat", ["/nonexistent", "0.5", "n", "n", "q", "0.3"]),   # bad threshold type
    ("Cp", ["/nonexistent", "-3", "gk"]),                            # negative random-network count
    ("CUBFW_Lp", ["/nonexistent", "-1", "gl"]), ("CUmodularity", ["/nonexistent", "-1"]),
    ("Maslov", []), ("Maslov", ["net.txt"]), ("Maslov", ["a.csr", "b.csr"]),
    ("SmallWorldProperty", []), ("SmallWorldProperty", ["a", "b"]),
])
def test_tools_reject_bad_arguments(tool, args):
    exe = os.path.join(BIN, tool)
    if not os.path.exists(exe):
        pytest.skip("tools not built")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)


def _mask_volume(path, values, dt, dims=(5, 4, 3)):
    formats.write_nifti(str(path), values, dt)
    raw = bytearray(path.read_bytes())
    h = np.frombuffer(bytes(raw[:352]), np.int16).copy()
    h[21], h[22], h[23] = dims
    h[40] = 1234  # an unrelated header word the conversion must carry over
    raw[:352] = h.tobytes()
    path.write_bytes(bytes(raw))


@pytest.mark.parametrize("dt,thr", [(formats.DT_UINT8, None), (formats.DT_FLOAT32, "0.7")])
def test_convertnii_tool(tmp_path, dt, thr):
    exe = os.path.join(BIN, "ConvertNII")
    if not os.path.exists(exe):
        pytest.skip("tools not built")
    rng = np.random.default_rng(5)
    if dt == formats.DT_UINT8:
        m = (rng.random(60) < 0.6).astype(np.uint8)
        sel = m >= 1
    else:
        # 0.7f is below the double 0.7: the reference's double cut drops it
        m = rng.choice(np.array([0.0, 0.5, 0.7, 0.9, 1.0], np.float32), 60)
        sel = m.astype(np.float64) >= 0.7
        assert (m == np.float32(0.7)).any()
    d = tmp_path / "res"
    d.mkdir()
    _mask_volume(tmp_path / "mask.nii", m, dt)
    x = rng.standard_normal(int(sel.sum())).astype(np.float32)
    formats.write_nm(str(d / "sub01_deg.nm"), x)
    formats.write_nm(str(d / "sub02_deg.nm"), x[:-1])  # wrong length: skipped
    formats.write_nm(str(d / "notes.nmx"), x)          # not a .nm file
    args = [str(d), str(tmp_path / "mask.nii")] + ([thr] if thr else [])
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "doesn't match" in r.stdout
    assert not (d / "sub02_deg.nii").exists() and not (d / "notes.nii").exists()
    out = (d / "sub01_deg.nii").read_bytes()
    mh = np.frombuffer((tmp_path / "mask.nii").read_bytes()[:352], np.int16)
    oh = np.frombuffer(out[:352], np.int16)
    diff = np.nonzero(mh != oh)[0].tolist()
    assert set(diff) <= {35, 36} and oh[35] == 16 and oh[36] == 32 and oh[40] == 1234
    vol = np.frombuffer(out[352:], np.float32)
    assert vol.size == 60
    np.testing.assert_array_equal(vol[sel], x)
    assert np.all(vol[~sel] == 0)
    # the Python mirror writes the same bytes
    assert formats.nm_to_nii(str(d / "sub01_deg.nm"), str(tmp_path / "mask.nii"), str(tmp_path / "py.nii"),
                             float(thr) if thr else 0.5)
    assert (tmp_path / "py.nii").read_bytes() == out
    assert not formats.nm_to_nii(str(d / "sub02_deg.nm"), str(tmp_path / "mask.nii"), str(tmp_path / "q.nii"))


def test_small_world_property_tool(tmp_path):
    exe = os.path.join(BIN, "SmallWorldProperty")
    if not os.path.exists(exe):
        pytest.skip("tools not built")
    files = {"sub01_1317_spa1.000%_cor0.300_lambda.txt": "1.25", "sub01_1317_spa1.000%_cor0.300_gamma.txt": "2.5",
             "sub02_1317_spa2.000%_cor0.250_lambda.txt": "1.1", "sub02_1317_spa2.000%_cor0.250_gamma.txt": "3.3",
             "sub03_1317_spa5.000%_cor0.100_lambda.txt": "1.0"}  # no gamma: left alone
    for name, v in files.items():
        (tmp_path / name).write_text(v + "\n")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.stdout, r.stderr)
    sw1 = float((tmp_path / "sub01_1317_spa1.000%_cor0.300_smallWorld.txt").read_text())
    sw2 = float((tmp_path / "sub02_1317_spa2.000%_cor0.250_smallWorld.txt").read_text())
    assert sw1 == 2.0
    assert abs(sw2 - 3.3 / 1.1) < 5e-6 * 3.0  # six significant digits
    assert not (tmp_path / "sub03_1317_spa5.000%_cor0.100_smallWorld.txt").exists()
    assert "no matching gamma" in r.stdout
