"""LMFX: the fixture file format of this repository's golden vectors.

A small binary container that C reads as easily as Python (tests/golden/
lmfx.h): libmat's own storage, so a C program can point a Mat or a Perm
straight at an entry's payload.  Little-endian:

    magic      8 bytes  "LMATFX01"
    count      u32, then 4 bytes of zeros
    count entries, each:
      name_len u16, name (ASCII, name_len bytes), zeros up to a multiple
               of 8 bytes from the start of the file
      kind     u8    1 Mat f32, 2 Mat f64, 3 Perm (size_t = u64),
                     4 int64 scalar, 5 f64 scalar
      rank     u8    dimensions of the array written (0, 1 or 2)
      pad      u8 x 6
      rows     u64, cols u64
      payload  rows*cols elements, column-major (Mat: data[j*rows + i],
               mat.h:387-393; Perm: its data array, mat.h:382-385),
               padded with zeros to a multiple of 8 bytes

A 1-D float array is stored as a Vec (rows x 1 Mat, mat.h:379, rank 1); a
1-D integer array as a Perm; 0-d values as scalars.  Reading gives back the
numpy arrays that were written (natural (row, col) indexing).
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"LMATFX01"
MAT_F32, MAT_F64, PERM, I64, F64 = 1, 2, 3, 4, 5


def _entry(name: str, v) -> bytes:
    a = np.asarray(v)
    if a.ndim == 0:
        if np.issubdtype(a.dtype, np.integer):
            kind, rows, cols, payload = I64, 1, 1, np.int64(a).tobytes()
        else:
            kind, rows, cols, payload = F64, 1, 1, np.float64(a).tobytes()
    elif np.issubdtype(a.dtype, np.integer):
        assert a.ndim == 1, name
        kind, rows, cols = PERM, a.shape[0], 1
        payload = a.astype("<u8").tobytes()
    else:
        kind = MAT_F32 if a.dtype == np.float32 else MAT_F64
        assert a.dtype in (np.float32, np.float64), (name, a.dtype)
        m = a.reshape(-1, 1) if a.ndim == 1 else a
        rows, cols = m.shape
        payload = np.asfortranarray(m).ravel(order="F").astype(a.dtype.newbyteorder("<")).tobytes()
    nb = name.encode("ascii")
    head = struct.pack("<H", len(nb)) + nb
    head += b"\0" * (-len(head) % 8)
    head += struct.pack("<BB6xQQ", kind, a.ndim, rows, cols)
    return head + payload + b"\0" * (-len(payload) % 8)


def write(path: str, entries: dict) -> None:
    """Entries in name order; every entry starts 8-byte aligned (the file
    header is 16 bytes: magic, count, 4 bytes of zeros)."""
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<I4x", len(entries)))
        for name in sorted(entries):
            f.write(_entry(name, entries[name]))


def read(path: str) -> dict:
    with open(path, "rb") as f:
        buf = f.read()
    assert buf[:8] == MAGIC, f"{path}: not an LMFX file"
    (count,) = struct.unpack_from("<I", buf, 8)
    off, out = 16, {}
    for _ in range(count):
        (nl,) = struct.unpack_from("<H", buf, off)
        name = buf[off + 2:off + 2 + nl].decode("ascii")
        off += 2 + nl
        off += -off % 8
        kind, rank, rows, cols = struct.unpack_from("<BB6xQQ", buf, off)
        off += 24
        n = rows * cols
        if kind in (MAT_F32, MAT_F64):
            dt = np.dtype("<f4" if kind == MAT_F32 else "<f8")
            a = np.frombuffer(buf, dt, n, off).reshape((rows, cols), order="F")
            val = a[:, 0].copy() if rank == 1 else a.copy()
            size = n * dt.itemsize
        elif kind == PERM:
            val = np.frombuffer(buf, "<u8", n, off).astype(np.int64)
            size = n * 8
        elif kind == I64:
            val = np.int64(np.frombuffer(buf, "<i8", 1, off)[0])
            size = 8
        elif kind == F64:
            val = np.float64(np.frombuffer(buf, "<f8", 1, off)[0])
            size = 8
        else:
            raise ValueError(f"{path}: entry {name}: unknown kind {kind}")
        off += size + (-size % 8)
        out[name] = val
    return out
