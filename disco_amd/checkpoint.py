"""Reader and writer of the checkpoint container (disco_amd/csrc/checkpoint.h).

A checkpoint holds the datasets of the reference's HDF5 checkpoint
(Output/h5out.c:366-622) by their HDF5 paths -- ``Grid/T``, ``Grid/r_jph``,
``Grid/Index``, ``Data/Cells`` ... -- as NumPy arrays of the same shapes, so
analysis code written against the HDF5 layout (``f["Data/Cells"][...]``)
works on the dict ``read`` returns.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"DISCOCKP"
VERSION = 1
_TYPES = {0: np.float64, 1: np.int32, 2: np.uint8}
_CODES = {np.dtype(np.float64): 0, np.dtype(np.int32): 1, np.dtype(np.uint8): 2}


def read(path: str) -> dict:
    """{dataset path: array}; byte datasets (the Opts strings) as str."""
    out = {}
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a checkpoint")
        ver, n = struct.unpack("<II", f.read(8))
        if ver != VERSION:
            raise ValueError(f"{path}: version {ver}, expected {VERSION}")
        for _ in range(n):
            (nl,) = struct.unpack("<I", f.read(4))
            name = f.read(nl).decode()
            ty, rank = struct.unpack("<II", f.read(8))
            dims = struct.unpack(f"<{rank}Q", f.read(8 * rank)) if rank else ()
            (nb,) = struct.unpack("<Q", f.read(8))
            buf = f.read(nb)
            if ty == 2:
                out[name] = buf.decode()
            else:
                out[name] = np.frombuffer(buf, dtype=_TYPES[ty]).reshape(dims).copy()
    return out


def write(path: str, sets: dict) -> None:
    """The inverse of read (arrays of float64 or int32, and str)."""
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<II", VERSION, len(sets)))
        for name, v in sets.items():
            if isinstance(v, str):
                a, ty = np.frombuffer(v.encode(), np.uint8), 2
            else:
                a = np.ascontiguousarray(v)
                ty = _CODES[a.dtype]
            nm = name.encode()
            f.write(struct.pack("<I", len(nm)) + nm + struct.pack("<II", ty, a.ndim))
            f.write(struct.pack(f"<{a.ndim}Q", *a.shape))
            f.write(struct.pack("<Q", a.nbytes) + a.tobytes())
