"""Reader of the correlator record file written by PLEGMA_Correlator::writeHDF5
(include/plegma/PLEGMA_h5twin.h), the on-disk twin of the reference's HDF5
correlator file (lib/PLEGMA_Correlator.cu:893-946) for a build without HDF5.

    file   := b"PLGH5TW1" record*
    record := u32 0x54455344 | u32 len, path | u32 dtype (1 f64, 2 f32, 3 i32)
              | u32 ndim, u64 dims[ndim] | u32 nattr, (u32 len, name, u32 len, value)*
              | u64 nbytes, elements (little-endian, C order)

read() returns {path: Dataset}; to_hdf5() writes the same objects into an HDF5
file with h5py (an optional dependency, absent from this image).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

MAGIC = b"PLGH5TW1"
TAG = 0x54455344
DTYPES = {1: np.dtype("<f8"), 2: np.dtype("<f4"), 3: np.dtype("<i4")}


@dataclass
class Dataset:
    data: np.ndarray
    attrs: dict = field(default_factory=dict)


def read(path: str) -> dict:
    """All datasets of the record file at `path`, in file order."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:8] != MAGIC:
        raise ValueError(f"{path}: not a PLGH5TW1 record file")
    out = {}
    o = 8

    def take(n):
        nonlocal o
        if o + n > len(buf):
            raise ValueError(f"{path}: truncated record")
        b = buf[o:o + n]
        o += n
        return b

    def u32():
        return struct.unpack("<I", take(4))[0]

    def string():
        return take(u32()).decode()

    while o < len(buf):
        if u32() != TAG:
            raise ValueError(f"{path}: corrupt record at byte {o - 4}")
        name = string()
        dtype = u32()
        if dtype not in DTYPES:
            raise ValueError(f"{path}: {name}: unknown element type {dtype}")
        ndim = u32()
        dims = struct.unpack(f"<{ndim}Q", take(8 * ndim))
        attrs = {}
        for _ in range(u32()):
            k = string()
            attrs[k] = string()
        nbytes = struct.unpack("<Q", take(8))[0]
        dt = DTYPES[dtype]
        if nbytes != int(np.prod(dims, dtype=np.int64)) * dt.itemsize:
            raise ValueError(f"{path}: {name}: {nbytes} bytes for shape {dims}")
        data = np.frombuffer(take(nbytes), dtype=dt).reshape(dims)
        out[name] = Dataset(data.copy(), attrs)
    return out


def to_hdf5(twin_path: str, h5_path: str | None = None) -> str:
    """Writes the objects of a record file into an HDF5 file (default: the
    record file's name without ".twin").  Needs h5py."""
    import h5py  # optional; not installed in this image

    h5_path = h5_path or (twin_path[:-5] if twin_path.endswith(".twin") else twin_path + ".h5")
    with h5py.File(h5_path, "a") as f:
        for name, ds in read(twin_path).items():
            if name in f:
                continue
            d = f.create_dataset(name, data=ds.data)
            for k, v in ds.attrs.items():
                d.attrs[k] = np.bytes_(v)
    return h5_path
