"""PAGANI file formats in Python (mirror of host/io.cpp).

    .nii  NIfTI-1, 352-byte header; the tools read it as int16 words: dims at
          words 21..24, datatype at 35, bitpix at 36; data from byte 352
          (U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:138-235)
    .csr  int32 N+1, int32 R[N+1], int32 nnz, int32 C[nnz] [, int32 nnz, float32 V]
    .nm   int32 N, float32[N]
"""
from __future__ import annotations

import struct

import numpy as np

HDR = 352
DT_UINT8, DT_FLOAT32, DT_FLOAT64 = 2, 16, 64


def write_nifti(path: str, data: np.ndarray, datatype: int) -> None:
    """data: [x, y, z] or [t, z, y, x]-ordered volume flattened as the tools
    index it (voxel index fastest, then time).  Shape (T, V) or (V,)."""
    data = np.asarray(data)
    if data.ndim == 1:
        t, v = 1, data.size
    else:
        t, v = data.shape
    hdr = np.zeros(HDR // 2, np.int16)
    hdr[0] = 348  # sizeof_hdr (low word)
    nx = v
    hdr[20] = 4 if t > 1 else 3
    hdr[21], hdr[22], hdr[23] = nx, 1, 1
    hdr[24] = t
    hdr[35] = datatype
    bitpix = {DT_UINT8: 8, DT_FLOAT32: 32, DT_FLOAT64: 64}[datatype]
    hdr[36] = bitpix
    raw = hdr.tobytes()
    raw = raw[:108] + struct.pack("<f", 352.0) + raw[112:]  # vox_offset
    dt = {DT_UINT8: np.uint8, DT_FLOAT32: np.float32, DT_FLOAT64: np.float64}[datatype]
    with open(path, "wb") as f:
        f.write(raw)
        f.write(np.ascontiguousarray(data, dtype=dt).tobytes())


def read_csr(path: str, weighted: bool | None = None):
    """(row, col, val-or-None); weighted=None detects trailing values."""
    buf = open(path, "rb").read()
    rlen = struct.unpack_from("<i", buf, 0)[0]
    row = np.frombuffer(buf, np.int32, rlen, 4)
    off = 4 + 4 * rlen
    clen = struct.unpack_from("<i", buf, off)[0]
    col = np.frombuffer(buf, np.int32, clen, off + 4)
    off += 4 + 4 * clen
    val = None
    has = len(buf) > off
    if weighted or (weighted is None and has):
        vlen = struct.unpack_from("<i", buf, off)[0]
        if vlen != clen:
            raise ValueError("csr damaged: Vlength != Clength")
        val = np.frombuffer(buf, np.float32, vlen, off + 4)
    return row.copy(), col.copy(), (None if val is None else val.copy())


def write_csr(path: str, row, col, val=None) -> None:
    row = np.asarray(row, np.int32)
    col = np.asarray(col, np.int32)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", row.size))
        f.write(row.tobytes())
        f.write(struct.pack("<i", col.size))
        f.write(col.tobytes())
        if val is not None:
            f.write(struct.pack("<i", col.size))
            f.write(np.asarray(val, np.float32).tobytes())


def read_nm(path: str) -> np.ndarray:
    buf = open(path, "rb").read()
    n = struct.unpack_from("<i", buf, 0)[0]
    return np.frombuffer(buf, np.float32, n, 4).copy()


def write_nm(path: str, x) -> None:
    x = np.asarray(x, np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", x.size))
        f.write(x.tobytes())


def csr_name(subject_base: str, n: int, s_thresh: float, r_thresh: float, weighted: bool = False) -> str:
    """U CorMat.cpp:284,333-335: <base>_<N>_spa%.3f%%_cor%.3f.csr."""
    return f"{subject_base}_{n}_spa{s_thresh:.3f}%_cor{r_thresh:.3f}" + ("_weighted.csr" if weighted else ".csr")


def nm_to_nii(nm_path: str, mask_path: str, out_path: str, thr: float = 0.5) -> bool:
    """One .nm vector back to a NIfTI volume through the mask, as the
    ConvertNII tool does (U/ConvertNII/ConvertNII/ConvertNII.cpp:107-170): the
    mask's 352-byte header with datatype 16 and bitpix 32, then float32 values
    at voxels with (double) mask >= thr, 0 elsewhere.  Returns False (and
    writes nothing) when the vector length differs from the mask's count."""
    raw = open(mask_path, "rb").read()
    hdr = np.frombuffer(raw, np.int16, HDR // 2).copy()
    total = int(hdr[21]) * int(hdr[22]) * int(hdr[23])
    if hdr[35] == DT_UINT8:
        mask = np.frombuffer(raw, np.uint8, total, HDR).astype(np.float32)
    elif hdr[35] == DT_FLOAT32:
        mask = np.frombuffer(raw, np.float32, total, HDR)
    else:
        raise ValueError("mask data-type error, only unsigned char and float can be handled")
    sel = mask.astype(np.float64) >= float(thr)
    x = read_nm(nm_path)
    if x.size != int(sel.sum()):
        return False
    vol = np.zeros(total, np.float32)
    vol[sel] = x
    hdr[35], hdr[36] = DT_FLOAT32, 32
    with open(out_path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(vol.tobytes())
    return True
