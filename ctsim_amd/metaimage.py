"""MetaImage (.mhd header + .raw data) files, the cache format of CTSim's
util/io/mhd (writeMetaImage_mhd.m:26-129, readMetaImage_mhd.m, readMetaImage_raw.m).

Data are stored in MATLAB order (first dimension fastest), little endian.  The header
fields are written in the reference's order (NDims, DimSize, BinaryDataByteOrderMSB,
ElementSpacing, Position, ElementType, ElementByteOrderMSB, ElementDataFile, then any
other tags); the reader takes the data from the last prod(DimSize) elements of the
data file, as readMetaImage_raw.m does (so headers embedded in front are skipped)."""
from __future__ import annotations

import os

import numpy as np

_TYPES = {
    "MET_UCHAR": np.uint8, "MET_CHAR": np.int8, "MET_SHORT": np.int16, "MET_USHORT": np.uint16,
    "MET_FLOAT": np.float32, "MET_DOUBLE": np.float64,
}
_FROM_NP = {np.dtype(v): k for k, v in _TYPES.items()}


def write_metaimage(filename: str, data: np.ndarray, spacing=None, origin=None, extra: dict | None = None) -> str:
    """Write data (any of the reference's element types; other types are stored as
    MET_FLOAT as writeMetaImage.m does) to name.mhd + name.raw; returns the .mhd path."""
    base, ext = os.path.splitext(filename)
    if ext.lower() not in ("", ".mhd"):
        raise ValueError("The given extension is not supported!")
    mhd, raw = base + ".mhd", base + ".raw"
    a = np.asarray(data)
    if a.dtype not in _FROM_NP:
        a = a.astype(np.float32)
    etype = _FROM_NP[a.dtype]
    dims = [1] if a.ndim == 0 else list(a.shape)
    if a.ndim == 2 and 1 in a.shape:  # MATLAB vectors are 1-D
        dims = [a.size]
    n = len(dims)
    spacing = np.ones(n) if spacing is None else np.asarray(spacing, float).ravel()
    if spacing.size != n or spacing.min() < 0:
        raise ValueError("You have to specify a correct ElementSpacing of the data!")
    lines = [f"NDims = {n}", "DimSize = " + " ".join(str(int(d)) for d in dims),
             "BinaryDataByteOrderMSB = False", "ElementSpacing = " + " ".join(f"{s:f}" for s in spacing)]
    if origin is not None:
        origin = np.asarray(origin, float).ravel()
        if origin.size != n:
            raise ValueError("You have to specify a correct Position of the data!")
        lines.append("Position = " + " ".join(f"{p:f}" for p in origin))
    lines += [f"ElementType = {etype}", "ElementByteOrderMSB = False",
              f"ElementDataFile = {os.path.basename(raw)}"]
    for k, v in (extra or {}).items():
        lines.append(f"{k} = {v}")
    with open(mhd, "w") as f:
        f.write("\n".join(lines) + "\n")
    np.asarray(a, a.dtype.newbyteorder("<")).ravel(order="F").tofile(raw)
    return mhd


def read_metaimage(filename: str):
    """Read name.mhd (+ its ElementDataFile); returns (data in MATLAB order, meta dict)."""
    if not filename.lower().endswith(".mhd"):
        raise ValueError(f'The given extension (of file "{filename}") is not supported!')
    meta = {}
    with open(filename) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            name, eq, value = line.partition("=")
            if not eq or not name.strip() or not value.strip():
                raise ValueError("File parsing error!")
            meta[name.strip()] = value.strip()
    n = int(meta["NDims"])
    dims = [int(x) for x in meta["DimSize"].split()][:n]
    if len(dims) != n or min(dims) <= 0:
        raise ValueError("Illegal DimSize!")
    etype = meta.get("ElementType")
    if etype not in _TYPES:
        raise ValueError("Unknown ElementType!")
    for key in ("BinaryDataByteOrderMSB", "ElementByteOrderMSB"):
        if meta.get(key, "False") == "True":
            raise ValueError("Specified byteOrderMSB is not implemented!")
    raw = meta["ElementDataFile"]
    if not os.path.isabs(raw):
        raw = os.path.join(os.path.dirname(os.path.abspath(filename)), raw)
    dt = np.dtype(_TYPES[etype]).newbyteorder("<")
    count = int(np.prod(dims))
    with open(raw, "rb") as f:
        f.seek(-count * dt.itemsize, os.SEEK_END)  # the data are the file's last bytes
        data = np.fromfile(f, dt, count)
    out = {"NDims": n, "DimSize": dims, "ElementType": etype,
           "ElementSpacing": [float(x) for x in meta.get("ElementSpacing", " ".join(["1"] * n)).split()]}
    if "Position" in meta:
        out["Position"] = [float(x) for x in meta["Position"].split()]
    for k, v in meta.items():
        out.setdefault(k, v)
    return data.astype(dt.newbyteorder("=")).reshape(dims, order="F") if n > 1 else data.astype(
        dt.newbyteorder("=")), out
