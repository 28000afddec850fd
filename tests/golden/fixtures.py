"""Reader for the BMFX fixtures that make_fixtures.mjs writes from the reference.

A file is gzip of: b"BMFX0001", a little-endian u32 header length, a JSON header
({"arrays": [{name, dtype, length, offset}], "meta": {...}}), zero padding to 8 bytes,
then each array's raw little-endian bytes at its offset (each 8-byte aligned).
"""
import glob
import gzip
import json
import os

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
_DT = {"f64": np.float64, "f32": np.float32, "i32": np.int32, "u32": np.uint32, "u8": np.uint8, "i8": np.int8}


class Fixture(dict):
    """name -> numpy array, with the header's meta dict in .meta."""

    meta: dict


def load(path):
    raw = gzip.open(path, "rb").read()
    if raw[:8] != b"BMFX0001":
        raise ValueError(f"{path}: not a BMFX fixture")
    n = int.from_bytes(raw[8:12], "little")
    header = json.loads(raw[12:12 + n].decode())
    base = 12 + n + ((8 - ((12 + n) % 8)) % 8)
    out = Fixture()
    for e in header["arrays"]:
        dt = np.dtype(_DT[e["dtype"]])
        start = base + e["offset"]
        out[e["name"]] = np.frombuffer(raw, dtype=dt, count=e["length"], offset=start).copy()
    out.meta = header["meta"]
    return out


def find(kind, scene=None):
    """Paths of fixtures of a kind ('soft', 'trace', 'colour', 'traj'), sorted."""
    pat = f"{scene or '*'}.{kind}.*bmfx.gz" if kind != "traj" else f"{scene or '*'}.traj.bmfx.gz"
    return sorted(glob.glob(os.path.join(DATA, pat)))
