"""ORACLE — test infrastructure only. ctypes binding of _build/liboracle.so (oracle.h)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I32 = ctypes.POINTER(ctypes.c_int32)
_U32 = ctypes.POINTER(ctypes.c_uint32)

# (struct field, kind, fixture column, numbers per unit, unit) in oracle.h order.
BODY = [
    ("position", _D, "position", 3), ("rotation", _D, "rotation", 4), ("velocity", _D, "velocity", 3),
    ("angular", _D, "angular", 3), ("dims", _D, "dims", 3), ("inv_mass", _D, "invMass", 1),
    ("inertia", _D, "inertia", 3), ("inv_inertia", _D, "invInertia", 3), ("inv_i", _D, "invI", 6),
    ("axes", _D, "axes", 9), ("tumbled", _D, "tumbled", 3), ("ceiling", _D, "ceiling", 1),
    ("gravity_scale", _D, "gravityScale", 1), ("damping", _D, "damping", 1),
    ("angular_damping", _D, "angularDamping", 1), ("start", _D, "start", 3), ("start_q", _D, "startQ", 4),
    ("type", _U8, "type", 1), ("shape", _U8, "shape", 1), ("frozen", _U8, "frozen", 1), ("lock", _U8, "lock", 1),
    ("framed", _U8, "framed", 1), ("awake", _U8, "awake", 1), ("joint_head", _I32, "jointHead", 1),
]
MANIFOLD = [
    ("a", _I32, "a", 1), ("b", _I32, "b", 1), ("pair", _I32, "pair", 1), ("sub", _I32, "sub", 1),
    ("row", _I32, "row", 1), ("points", _U8, "points", 1), ("sliding", _U8, "sliding", 1),
    ("normal", _D, "normal", 3), ("t1", _D, "t1", 3), ("t2", _D, "t2", 3), ("friction", _D, "friction", 1),
    ("restitution", _D, "restitution", 1), ("tolerance", _D, "tolerance", 1), ("patch", _D, "patch", 1),
    ("roll_impulse1", _D, "rollImpulse1", 1), ("roll_impulse2", _D, "rollImpulse2", 1),
    ("spin_impulse", _D, "spinImpulse", 1), ("tangent_impulse1", _D, "tangentImpulse1", 1),
    ("tangent_impulse2", _D, "tangentImpulse2", 1), ("drift1", _D, "drift1", 1), ("drift2", _D, "drift2", 1),
    ("separation", _D, "separation", 4), ("point", _D, "point", 12), ("shift", _D, "shift", 12),
    ("normal_impulse", _D, "normalImpulse", 4), ("max_normal_impulse", _D, "maxNormalImpulse", 4),
    ("relative_velocity", _D, "relativeVelocity", 4), ("gap", _D, "gap", 1), ("now", _D, "now", 4),
    ("ra", _D, "ra", 12), ("rb", _D, "rb", 12), ("la", _D, "la", 12), ("lb", _D, "lb", 12),
    ("adjusted", _D, "adjusted", 4), ("normal_mass", _D, "normalMass", 4), ("jac", _D, "jac", 48),
    ("block", _D, "block", 16), ("lever", _D, "lever", 4), ("anchor_a", _D, "anchorA", 3),
    ("anchor_b", _D, "anchorB", 3), ("fjac", _D, "fjac", 24), ("tangent_mass", _D, "tangentMass", 3),
    ("roll_mass1", _D, "rollMass1", 1), ("roll_mass2", _D, "rollMass2", 1), ("spin_mass", _D, "spinMass", 1),
    ("pid", _U32, "pid", 4),
]
PAIR = [("pair_points", _U8, "points", 1), ("pair_pid", _U32, "pid", 4),
        ("pair_normal_impulse", _D, "normalImpulse", 4), ("pair_twist", _D, "twist", 3),
        ("pair_friction", _D, "friction", 3), ("pair_slip", _D, "slip", 3)]

_NP = {_D: np.float64, _U8: np.uint8, _I32: np.int32, _U32: np.uint32}


class OracleSoft(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int32) for n in ("top", "awake_count", "substeps", "count", "solved", "pair_count")]
                + [("dt", ctypes.c_double), ("gravity", ctypes.c_double * 3)]
                + [(f, k) for f, k, _, _ in BODY[:17]] + [(f, k) for f, k, _, _ in BODY[17:23]]
                + [("joint_head", _I32), ("awake_list", _I32)]
                + [("starts", _I32)] + [(f, k) for f, k, _, _ in MANIFOLD[:5]] + [(f, k) for f, k, _, _ in MANIFOLD[5:7]]
                + [(f, k) for f, k, _, _ in MANIFOLD[7:45]] + [("pid", _U32)]
                + [(f, k) for f, k, _, _ in PAIR])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: run make -C {HERE}")
        _lib = ctypes.CDLL(LIB)
        _lib.oracle_soft_step.argtypes = [ctypes.POINTER(OracleSoft)]
        _lib.oracle_soft_step.restype = ctypes.c_int
        _lib.oracle_soft_step_workers.argtypes = [ctypes.POINTER(OracleSoft), ctypes.c_int]
        _lib.oracle_soft_step_workers.restype = ctypes.c_int
    return _lib


def colour(a, b, top, type_, frozen, frozen_count):
    """oracle_colour: (colour, order, row, starts, solved) for manifolds a[m]-b[m]."""
    lib_ = lib()
    n = len(a)
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    type_ = np.ascontiguousarray(type_, np.uint8)
    frozen = np.ascontiguousarray(frozen, np.uint8)
    col, order, row = (np.zeros(max(n, 1), np.int32) for _ in range(3))
    starts = np.zeros(35, np.int32)
    solved = ctypes.c_int32()
    f = lib_.oracle_colour
    f.restype = ctypes.c_int
    f(n, a.ctypes.data_as(_I32), b.ctypes.data_as(_I32), top, type_.ctypes.data_as(_U8), frozen.ctypes.data_as(_U8),
      frozen_count, col.ctypes.data_as(_I32), order.ctypes.data_as(_I32), row.ctypes.data_as(_I32),
      starts.ctypes.data_as(_I32), ctypes.byref(solved))
    return col[:n], order[:n], row[:n], starts, solved.value


def columns(fx):
    """The soft step's columns from a fixture's 'in.*' arrays (tests/golden/fixtures.py
    soft_columns: bodies as ``b.<name>``, manifolds ``c.<name>``, pairs ``p.<name>``)."""
    import fixtures
    return fixtures.soft_columns(fx)


def soft_step(fx, cols=None, workers=0):
    """Runs oracle_soft_step on a fixture's input state (with `workers` more threads:
    oracle_soft_step_workers); returns the columns after it."""
    meta = fx.meta
    cols = columns(fx) if cols is None else cols
    st = OracleSoft()
    st.top, st.awake_count, st.substeps = meta["top"], meta["awake"], meta["substeps"]
    st.count, st.solved, st.pair_count = meta["count"], meta["solved"], meta["pairs"]
    st.dt = meta["dt"]
    for k in range(3):
        st.gravity[k] = cols["gravity"][k]

    def ptr(a, kind):
        return a.ctypes.data_as(kind)

    for f, kind, name, _ in BODY:
        setattr(st, f, ptr(cols["b." + name], kind))
    for f, kind, name, _ in MANIFOLD:
        setattr(st, f, ptr(cols["c." + name], kind))
    for f, kind, name, _ in PAIR:
        setattr(st, f, ptr(cols["p." + name], kind))
    st.awake_list = ptr(cols["awake_list"], _I32)
    st.starts = ptr(cols["starts"], _I32)
    rc = (lib().oracle_soft_step_workers(ctypes.byref(st), workers) if workers
          else lib().oracle_soft_step(ctypes.byref(st)))
    if rc != 0:
        raise RuntimeError(f"oracle_soft_step returned {rc}")
    return cols


def collide(shapes_a, shapes_b, margin, tolerance):
    """oracle_collide on flat shape records (kind, centre 3, axes 9, half 3, weight 3):
    (count, normal, points, separations, ids) per pair."""
    n = len(margin)
    a = np.ascontiguousarray(shapes_a, np.float64)
    b = np.ascontiguousarray(shapes_b, np.float64)
    margin = np.ascontiguousarray(margin, np.float64)
    tolerance = np.ascontiguousarray(tolerance, np.float64)
    count = np.zeros(n, np.int32)
    normal = np.zeros(3 * n)
    points = np.zeros(12 * n)
    separations = np.zeros(4 * n)
    ids = np.zeros(4 * n, np.uint32)
    f = lib().oracle_collide
    f.restype = ctypes.c_int
    rc = f(n, a.ctypes.data_as(_D), b.ctypes.data_as(_D), margin.ctypes.data_as(_D), tolerance.ctypes.data_as(_D),
           count.ctypes.data_as(_I32), normal.ctypes.data_as(_D), points.ctypes.data_as(_D),
           separations.ctypes.data_as(_D), ids.ctypes.data_as(_U32))
    if rc != n:
        raise ValueError(f"oracle_collide: pair {-rc - 1} is of a kind not restated")
    return count, normal, points, separations, ids


def _f64(x):
    return np.ascontiguousarray(x, np.float64)


def closest(shapes_a, shapes_b):
    """oracle_closest on shape records: (distance, nearest point on a, on b) per pair."""
    a, b = _f64(shapes_a), _f64(shapes_b)
    n = a.size // 19
    distance, near_a, near_b = np.zeros(n), np.zeros(3 * n), np.zeros(3 * n)
    lib().oracle_closest(n, a.ctypes.data_as(_D), b.ctypes.data_as(_D), distance.ctypes.data_as(_D),
                         near_a.ctypes.data_as(_D), near_b.ctypes.data_as(_D))
    return distance, near_a, near_b


def impact(bodies_a, bodies_b, dt, tolerance):
    """oracle_impact on body records (32 numbers each): (t, posed a, posed b) per pair,
    the posed shapes as centre 3 and axes 9."""
    a, b, dt, tolerance = _f64(bodies_a), _f64(bodies_b), _f64(dt), _f64(tolerance)
    n = len(dt)
    t, pa, pb = np.zeros(n), np.zeros(12 * n), np.zeros(12 * n)
    lib().oracle_impact(n, a.ctypes.data_as(_D), b.ctypes.data_as(_D), dt.ctypes.data_as(_D),
                        tolerance.ctypes.data_as(_D), t.ctypes.data_as(_D), pa.ctypes.data_as(_D),
                        pb.ctypes.data_as(_D))
    return t, pa, pb


def meet(bodies_a, bodies_b, pair):
    """oracle_meet on body records and pair scalars (margin, tolerance, dt, gravity 3):
    (count, normal, points, separations, ids, shift, flags) per pair."""
    a, b, pair = _f64(bodies_a), _f64(bodies_b), _f64(pair)
    n = pair.size // 6
    count, flags = np.zeros(n, np.int32), np.zeros(n, np.int32)
    normal, points, separations = np.zeros(3 * n), np.zeros(12 * n), np.zeros(4 * n)
    ids, shift, loss = np.zeros(4 * n, np.uint32), np.zeros(12 * n), np.zeros(n)
    f = lib().oracle_meet
    f.restype = ctypes.c_int
    rc = f(n, a.ctypes.data_as(_D), b.ctypes.data_as(_D), pair.ctypes.data_as(_D), count.ctypes.data_as(_I32),
           normal.ctypes.data_as(_D), points.ctypes.data_as(_D), separations.ctypes.data_as(_D),
           ids.ctypes.data_as(_U32), shift.ctypes.data_as(_D), flags.ctypes.data_as(_I32), loss.ctypes.data_as(_D))
    if rc != n:
        raise ValueError(f"oracle_meet: pair {-rc - 1} is of a kind not restated")
    return count, normal, points, separations, ids, shift, flags, loss


def sincos(x):
    """oracle_sincos: the reference's sine and cosine of each x."""
    x = _f64(x)
    s, c = np.zeros(x.size), np.zeros(x.size)
    lib().oracle_sincos(x.size, x.ctypes.data_as(_D), s.ctypes.data_as(_D), c.ctypes.data_as(_D))
    return s, c


def hulls(fx):
    """oracle_hulls with a fixture's hull table (hullV, hullF, hullL, vertices, start, loops, planes)."""
    i32 = lambda x: np.ascontiguousarray(x, np.int32)
    hv, hf, hl = i32(fx["hullV"]), i32(fx["hullF"]), i32(fx["hullL"])
    v, st, lp, pl = _f64(fx["vertices"]), i32(fx["start"]), i32(fx["loops"]), _f64(fx["planes"])
    lib().oracle_hulls(len(hv) - 1, hv.ctypes.data_as(_I32), hf.ctypes.data_as(_I32), hl.ctypes.data_as(_I32),
                       v.ctypes.data_as(_D), st.ctypes.data_as(_I32), lp.ctypes.data_as(_I32), pl.ctypes.data_as(_D))


def losses(n):
    """The last collide() call's Manifold.loss per pair."""
    out = np.zeros(n)
    lib().oracle_losses(out.ctypes.data_as(_D))
    return out


def fields(columns, rows, spacing, heights):
    """oracle_fields: heightfields by number, field f with columns[f] x rows[f] samples
    spacing[f] apart, heights[f] its samples row by row."""
    n = len(columns)
    cols, rws = np.ascontiguousarray(columns, np.int32), np.ascontiguousarray(rows, np.int32)
    sp = _f64(spacing)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(h) for h in heights])
    flat = _f64(np.concatenate([np.asarray(h, np.float64) for h in heights]) if n else np.zeros(0))
    lib().oracle_fields(n, cols.ctypes.data_as(_I32), rws.ctypes.data_as(_I32), sp.ctypes.data_as(_D),
                        off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), flat.ctypes.data_as(_D))


def field_edges(f, count):
    """oracle_field_edges: field f's triangles' edge flags (count) and folds (6 each)."""
    flags = np.zeros(count, np.int32)
    folds = np.zeros(6 * count)
    lib().oracle_field_edges(f, count, flags.ctypes.data_as(_I32), folds.ctypes.data_as(_D))
    return flags, folds


def level(field, position, axes, shapes, boxes, margin, tolerance, cap):
    """oracle_level (touchLevel) for each shape record against its field: manifolds per
    shape, then count, normal, points, separations, ids in cap slots a shape."""
    n = len(margin)
    field = np.ascontiguousarray(field, np.int32)
    position, axes, shapes, boxes = _f64(position), _f64(axes), _f64(shapes), _f64(boxes)
    margin, tolerance = _f64(margin), _f64(tolerance)
    manifolds = np.zeros(n, np.int32)
    count = np.zeros(cap * n, np.int32)
    normal = np.zeros(3 * cap * n)
    points = np.zeros(12 * cap * n)
    seps = np.zeros(4 * cap * n)
    ids = np.zeros(4 * cap * n, np.uint32)
    p = lambda x, t: x.ctypes.data_as(t)
    rc = lib().oracle_level(n, p(field, _I32), p(position, _D), p(axes, _D), p(shapes, _D), p(boxes, _D), p(margin, _D),
                            p(tolerance, _D), cap, p(manifolds, _I32), p(count, _I32), p(normal, _D), p(points, _D),
                            p(seps, _D), p(ids, ctypes.POINTER(ctypes.c_uint32)))
    if rc != n:
        raise RuntimeError(f"oracle_level: shape {-rc - 1} gave more than {cap} manifolds")
    return manifolds, count, normal, points, seps, ids
