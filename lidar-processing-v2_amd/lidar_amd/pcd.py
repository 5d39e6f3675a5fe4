"""PCD v0.7 reader (the format of the reference's data/*.pcd scans, loaded
there with pcl::io::loadPCDFile<pcl::PointXYZI>, dataloader.cpp:153-170).

Supports DATA ascii and binary with F/I/U fields of any size and count 1;
returns a structured numpy array whose field names are the file's.
"""
import numpy as np

_TYPES = {("F", 4): "<f4", ("F", 8): "<f8", ("I", 1): "i1", ("I", 2): "<i2", ("I", 4): "<i4", ("I", 8): "<i8",
          ("U", 1): "u1", ("U", 2): "<u2", ("U", 4): "<u4", ("U", 8): "<u8"}


def read_pcd(path):
    with open(path, "rb") as fh:
        raw = fh.read()
    header = {}
    pos = 0
    while True:
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode("ascii").strip()
        pos = end + 1
        if not line or line.startswith("#"):
            continue
        key, *vals = line.split()
        header[key.upper()] = vals
        if key.upper() == "DATA":
            break
    names = header["FIELDS"]
    sizes = [int(v) for v in header["SIZE"]]
    types = header["TYPE"]
    counts = [int(v) for v in header.get("COUNT", ["1"] * len(names))]
    if any(c != 1 for c in counts):
        raise ValueError("fields with COUNT > 1 are not supported")
    n = int(header["POINTS"][0])
    dtype = np.dtype([(nm, _TYPES[(t, s)]) for nm, t, s in zip(names, types, sizes)])
    mode = header["DATA"][0]
    if mode == "binary":
        return np.frombuffer(raw, dtype, n, pos).copy()
    if mode == "ascii":
        rows = np.loadtxt(raw[pos:].decode("ascii").splitlines(), ndmin=2)[:n]
        out = np.zeros(n, dtype)
        for k, nm in enumerate(names):
            out[nm] = rows[:, k]
        return out
    raise ValueError(f"unsupported PCD DATA mode {mode!r}")
