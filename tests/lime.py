"""Minimal LIME / ILDG / SciDAC reader for the reference's data fixtures.

LIME record header (big-endian): magic 0x456789ab (u32), version (u16),
flags (u16), data length (u64), type (128 bytes, NUL padded); data follows,
padded to a multiple of 8 bytes.  ILDG gauge data: per site (t slowest, then z,
y, x fastest) four su3 in direction order x, y, z, t, big-endian doubles; nissa
reorders directions so that mu=0 is t (src/new_types/su3_op.hpp:217-225,
src/io/ILDG_File.hpp:111-125 with scidacMapping {0,3,2,1}).  SciDAC/ETMC
DiracFermion data: per site (same site order) spin-major spincolor.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = 0x456789AB


def records(path):
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos + 144 <= len(data):
        magic, ver, flags, length = struct.unpack(">IHHQ", data[pos:pos + 16])
        if magic != MAGIC:
            raise ValueError(f"bad LIME magic at {pos}")
        rtype = data[pos + 16:pos + 144].split(b"\0", 1)[0].decode()
        start = pos + 144
        out.append((rtype, data[start:start + length]))
        pos = start + ((length + 7) // 8) * 8
    return out


def record(path, rtype):
    for t, d in records(path):
        if t == rtype:
            return d
    raise KeyError(rtype)


def _file_to_lx(arr, L):
    """arr indexed by file order (t, z, y, x) -> nissa lx order (t, x, y, z)."""
    T, X, Y, Z = L
    a = arr.reshape((T, Z, Y, X) + arr.shape[1:])
    a = np.transpose(a, (0, 3, 2, 1) + tuple(range(4, a.ndim)))
    return np.ascontiguousarray(a.reshape((T * X * Y * Z,) + arr.shape[1:]))


def read_ildg_conf(path, L):
    """Return an lx quad_su3 conf (V,4,3,3,2) in nissa's direction order."""
    d = record(path, "ildg-binary-data")
    V = int(np.prod(L))
    a = np.frombuffer(d, dtype=">f8").astype(np.float64).reshape(V, 4, 3, 3, 2)
    a = _file_to_lx(a, L)
    out = np.empty_like(a)
    for mu in range(4):  # ILDG x,y,z,t -> nissa t,x,y,z
        out[:, (mu + 1) % 4] = a[:, mu]
    return out


def read_scidac_spincolor(path, L):
    d = record(path, "scidac-binary-data")
    V = int(np.prod(L))
    a = np.frombuffer(d, dtype=">f8").astype(np.float64).reshape(V, 4, 3, 2)
    return _file_to_lx(a, L)


def text_record(path, rtype):
    return record(path, rtype).decode(errors="replace")
