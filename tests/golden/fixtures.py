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
    out.path = path
    return out


def events_of(fx):
    """The reference's contact events along trajectory fixture `fx` (make_events.mjs), or None
    where none were written."""
    path = getattr(fx, "path", "").replace(".traj.bmfx.gz", ".events.bmfx.gz")
    return load(path) if path.endswith(".events.bmfx.gz") and os.path.exists(path) else None


def find(kind, scene=None):
    """Paths of fixtures of a kind ('soft', 'trace', 'colour', 'traj'), sorted."""
    pat = f"{scene or '*'}.{kind}.*bmfx.gz" if kind != "traj" else f"{scene or '*'}.traj.bmfx.gz"
    return sorted(glob.glob(os.path.join(DATA, pat)))


# The soft-step fixtures' columns ('in.*'): (fixture column, dtype, numbers a unit), for the
# bodies, the manifolds and the pair list (the reference's own names; src/world.js COLUMNS,
# src/contact.js Contacts, src/pairs.js).
SOFT_BODY = [("position", np.float64, 3), ("rotation", np.float64, 4), ("velocity", np.float64, 3),
             ("angular", np.float64, 3), ("dims", np.float64, 3), ("invMass", np.float64, 1),
             ("inertia", np.float64, 3), ("invInertia", np.float64, 3), ("invI", np.float64, 6),
             ("axes", np.float64, 9), ("tumbled", np.float64, 3), ("ceiling", np.float64, 1),
             ("gravityScale", np.float64, 1), ("damping", np.float64, 1), ("angularDamping", np.float64, 1),
             ("start", np.float64, 3), ("startQ", np.float64, 4), ("type", np.uint8, 1), ("shape", np.uint8, 1),
             ("frozen", np.uint8, 1), ("lock", np.uint8, 1), ("framed", np.uint8, 1), ("awake", np.uint8, 1),
             ("jointHead", np.int32, 1)]
SOFT_MANIFOLD = [("a", np.int32, 1), ("b", np.int32, 1), ("pair", np.int32, 1), ("sub", np.int32, 1),
                 ("row", np.int32, 1), ("points", np.uint8, 1), ("sliding", np.uint8, 1), ("normal", np.float64, 3),
                 ("t1", np.float64, 3), ("t2", np.float64, 3), ("friction", np.float64, 1),
                 ("restitution", np.float64, 1), ("tolerance", np.float64, 1), ("patch", np.float64, 1),
                 ("rollImpulse1", np.float64, 1), ("rollImpulse2", np.float64, 1), ("spinImpulse", np.float64, 1),
                 ("tangentImpulse1", np.float64, 1), ("tangentImpulse2", np.float64, 1), ("drift1", np.float64, 1),
                 ("drift2", np.float64, 1), ("separation", np.float64, 4), ("point", np.float64, 12),
                 ("shift", np.float64, 12), ("normalImpulse", np.float64, 4), ("maxNormalImpulse", np.float64, 4),
                 ("relativeVelocity", np.float64, 4), ("gap", np.float64, 1), ("now", np.float64, 4),
                 ("ra", np.float64, 12), ("rb", np.float64, 12), ("la", np.float64, 12), ("lb", np.float64, 12),
                 ("adjusted", np.float64, 4), ("normalMass", np.float64, 4), ("jac", np.float64, 48),
                 ("block", np.float64, 16), ("lever", np.float64, 4), ("anchorA", np.float64, 3),
                 ("anchorB", np.float64, 3), ("fjac", np.float64, 24), ("tangentMass", np.float64, 3),
                 ("rollMass1", np.float64, 1), ("rollMass2", np.float64, 1), ("spinMass", np.float64, 1),
                 ("pid", np.uint32, 4)]
SOFT_PAIR = [("points", np.uint8, 1), ("pid", np.uint32, 4), ("normalImpulse", np.float64, 4),
             ("twist", np.float64, 3), ("friction", np.float64, 3), ("slip", np.float64, 3)]


def soft_columns(fx):
    """A soft-step fixture's input columns as fresh arrays by fixture column name: bodies as
    ``b.<name>``, manifolds ``c.<name>``, pairs ``p.<name>`` (zeros where the fixture has none),
    and the awake list, colour starts and gravity."""
    meta = fx.meta
    top, count, pairs = meta["top"], meta["count"], meta["pairs"]
    cols = {}
    for name, dtype, per in SOFT_BODY:
        src = fx.get("in." + name)
        cols["b." + name] = src.copy() if src is not None else np.zeros(top * per, dtype)
    for name, dtype, per in SOFT_MANIFOLD:
        src = fx.get("in.c." + name)
        cols["c." + name] = src.copy() if src is not None else np.zeros(count * per, dtype)
    for name, dtype, per in SOFT_PAIR:
        cols["p." + name] = np.zeros(max(pairs, 1) * per, dtype)
    cols["awake_list"] = fx["in.awakeList"].copy()
    cols["starts"] = fx["in.starts"].copy()
    cols["gravity"] = fx["in.gravity"].copy()
    return cols



_NAMES = {np.dtype(v): k for k, v in _DT.items()}


def save(path, arrays, meta):
    """Writes named numpy arrays and a meta dict as a BMFX fixture (the layout load() reads)."""
    entries, offset, blobs = [], 0, []
    for name, a in arrays.items():
        a = np.ascontiguousarray(a)
        entries.append({"name": name, "dtype": _NAMES[a.dtype], "length": int(a.size), "offset": offset})
        raw = a.tobytes()
        pad = (8 - len(raw) % 8) % 8
        blobs.append(raw + b"\0" * pad)
        offset += len(raw) + pad
    header = json.dumps({"arrays": entries, "meta": meta}).encode()
    lead = 12 + len(header)
    body = b"BMFX0001" + len(header).to_bytes(4, "little") + header + b"\0" * ((8 - lead % 8) % 8) + b"".join(blobs)
    with gzip.open(path, "wb", compresslevel=9) as f:
        f.write(body)
