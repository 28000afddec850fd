"""ctypes binding of the C ABI in include/bearing_mi355x.h.

``Batch`` packs one or more worlds' columns (the reference's own layout, by slot,
manifold, point slot and pair) into the flat arrays and offset tables the library
takes, runs the soft step on the device, and splits the results back per world.
There is no CPU fallback: without the built library, or without a GPU, every call
raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libbearing_mi355x.so")
STARTS = 35

_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I32 = ctypes.POINTER(ctypes.c_int32)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_NP = {_D: np.float64, _U8: np.uint8, _I32: np.int32, _U32: np.uint32}

# bm_columns after n_worlds, in header order: (field, kind, unit, numbers per unit).
# Units: W1 (worlds + 1), W, S (35 per world), W3, B (body), A (awake), M (manifold), P (pair).
FIELDS = [
    ("body_offset", _I32, "W1", 1), ("manifold_offset", _I32, "W1", 1), ("pair_offset", _I32, "W1", 1),
    ("awake_offset", _I32, "W1", 1), ("starts", _I32, "S", 1), ("solved", _I32, "W", 1), ("substeps", _I32, "W", 1),
    ("dt", _D, "W", 1), ("gravity", _D, "W3", 1),
    ("position", _D, "B", 3), ("rotation", _D, "B", 4), ("velocity", _D, "B", 3), ("angular", _D, "B", 3),
    ("start", _D, "B", 3), ("start_q", _D, "B", 4), ("inv_i", _D, "B", 6), ("tumbled", _D, "B", 3),
    ("ceiling", _D, "B", 1), ("dims", _D, "B", 3), ("inv_mass", _D, "B", 1), ("inertia", _D, "B", 3),
    ("inv_inertia", _D, "B", 3), ("axes", _D, "B", 9), ("gravity_scale", _D, "B", 1), ("damping", _D, "B", 1),
    ("angular_damping", _D, "B", 1), ("type", _U8, "B", 1), ("shape", _U8, "B", 1), ("frozen", _U8, "B", 1),
    ("lock", _U8, "B", 1), ("framed", _U8, "B", 1), ("awake", _U8, "B", 1), ("joint_head", _I32, "B", 1),
    ("awake_list", _I32, "A", 1),
    ("a", _I32, "M", 1), ("b", _I32, "M", 1), ("pair", _I32, "M", 1), ("sub", _I32, "M", 1), ("row", _I32, "M", 1),
    ("colour", _I32, "M", 1), ("order", _I32, "M", 1), ("points", _U8, "M", 1), ("sliding", _U8, "M", 1), ("normal", _D, "M", 3), ("t1", _D, "M", 3),
    ("t2", _D, "M", 3), ("friction", _D, "M", 1), ("restitution", _D, "M", 1), ("tolerance", _D, "M", 1),
    ("patch", _D, "M", 1), ("roll_impulse1", _D, "M", 1), ("roll_impulse2", _D, "M", 1),
    ("spin_impulse", _D, "M", 1), ("tangent_impulse1", _D, "M", 1), ("tangent_impulse2", _D, "M", 1),
    ("drift1", _D, "M", 1), ("drift2", _D, "M", 1), ("separation", _D, "M", 4), ("point", _D, "M", 12),
    ("shift", _D, "M", 12), ("normal_impulse", _D, "M", 4), ("max_normal_impulse", _D, "M", 4),
    ("relative_velocity", _D, "M", 4), ("gap", _D, "M", 1), ("now", _D, "M", 4), ("ra", _D, "M", 12),
    ("rb", _D, "M", 12), ("la", _D, "M", 12), ("lb", _D, "M", 12), ("adjusted", _D, "M", 4),
    ("normal_mass", _D, "M", 4), ("jac", _D, "M", 48), ("block", _D, "M", 16), ("lever", _D, "M", 4),
    ("anchor_a", _D, "M", 3), ("anchor_b", _D, "M", 3), ("fjac", _D, "M", 24), ("tangent_mass", _D, "M", 3),
    ("roll_mass1", _D, "M", 1), ("roll_mass2", _D, "M", 1), ("spin_mass", _D, "M", 1), ("pid", _U32, "M", 4),
    ("pair_points", _U8, "P", 1), ("pair_pid", _U32, "P", 4), ("pair_normal_impulse", _D, "P", 4),
    ("pair_twist", _D, "P", 3), ("pair_friction", _D, "P", 3), ("pair_slip", _D, "P", 3),
]
PER_WORLD_UNITS = ("B", "A", "M", "P")


class Columns(ctypes.Structure):
    _fields_ = [("n_worlds", ctypes.c_int32)] + [(f, k) for f, k, _, _ in FIELDS]


# bm_columns_tap: void (*)(void* user, const bm_columns* columns)
COLUMNS_TAP = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(Columns))


def copy_columns(cols):
    """One world's bm_columns (n_worlds 1) copied into arrays by field name, as Batch takes them."""
    counts = {"B": cols.body_offset[1], "A": cols.awake_offset[1], "M": cols.manifold_offset[1],
              "P": cols.pair_offset[1]}
    out = {}
    for f, kind, unit, per in FIELDS:
        if unit == "W1":
            continue
        n = {"W": 1, "W3": 3, "S": STARTS}.get(unit, counts.get(unit, 0) * per)
        p = getattr(cols, f)
        out[f] = np.ctypeslib.as_array(p, (n,)).copy() if n and p else np.zeros(n, _NP[kind])
    return out


EXPORTS = ["bm_create", "bm_destroy", "bm_last_error", "bm_upload", "bm_soft_step", "bm_download",
           "bm_soft_step_host", "bm_last_kernel_ms", "bm_set_launch", "bm_lds_limit", "bm_colour", "bm_collide_host", "bm_meet_host", "bm_set_colour_lds", "bm_set_hulls",
           "bm_set_fields", "bm_level_host", "bm_set_wide", "bm_set_device_only", "bm_set_body_version",
           "bm_level_noalloc", "bm_level_packed", "bm_meet_indexed", "bm_meet_packed", "bm_world_set_parallel", "bm_batch_set_launch", "bm_contact_phase", "bm_set_graph",
           "bm_graph_counts", "bm_batch_set_graph", "bm_batch_graph_counts", "bm_world_moved", "bm_world_tap", "bm_world_events_of", "bm_world_set_broadphase",
           "bm_broad_set_bodies", "bm_broad_set_state", "bm_broad_keep", "bm_broad_set_trees", "bm_broad_bounds",
           "bm_broad_query", "bm_broad_holds",
           "bm_world_create", "bm_world_destroy", "bm_world_error", "bm_world_add", "bm_world_step", "bm_world_read",
           "bm_world_set_hulls", "bm_world_set_fields", "bm_world_times", "bm_upload_worlds", "bm_download_worlds",
           "bm_batch_create", "bm_batch_destroy", "bm_batch_error", "bm_batch_world", "bm_batch_step",
           "bm_batch_times", "bm_world_record", "bm_world_recorded", "bm_adjoint_create", "bm_adjoint_destroy",
           "bm_adjoint_error", "bm_adjoint_size", "bm_adjoint_recorded", "bm_adjoint_forward", "bm_adjoint_backward",
           "bm_adjoint_times", "bm_held_load", "bm_held_forward", "bm_held_backward", "bm_held_times",
           "bm_adjoint_branches", "bm_held_branches", "bm_adjoint_held"]

_lib = None


def context_signatures(lib):
    """The signatures of a context's calls (bm_create .. bm_last_kernel_ms) on `lib`."""
    vp = ctypes.c_void_p
    lib.bm_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32]
    lib.bm_destroy.argtypes = [vp]
    lib.bm_last_error.argtypes = [vp]
    lib.bm_last_error.restype = ctypes.c_char_p
    for name in ("bm_upload", "bm_download"):
        getattr(lib, name).argtypes = [vp, ctypes.POINTER(Columns)]
    lib.bm_soft_step.argtypes = [vp]
    lib.bm_colour.argtypes = [vp]
    lib.bm_last_kernel_ms.argtypes = [vp]
    lib.bm_last_kernel_ms.restype = ctypes.c_double


def load():
    """The built library, with its signatures set; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"bearing_amd: {LIB} is not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        context_signatures(lib)
        lib.bm_soft_step_host.argtypes = [vp, ctypes.POINTER(Columns)]
        lib.bm_collide_host.argtypes = [vp, ctypes.c_int32, _D, _D, _D, _D, _I32, _D, _D, _D, _U32, _D]
        lib.bm_meet_host.argtypes = [vp, ctypes.c_int32, _D, _D, _D, _I32, _D, _D, _D, _U32, _D, _I32, _D]
        lib.bm_set_hulls.argtypes = [vp, ctypes.c_int32, _I32, _I32, _I32, _D, _I32, _I32, _D]
        lib.bm_set_fields.argtypes = [vp, ctypes.c_int32, _I32, _I32, _D, ctypes.POINTER(ctypes.c_int64), _D]
        lib.bm_level_host.argtypes = [vp, ctypes.c_int32, _I32, _D, _D, _D, _D, _D, _D, ctypes.c_int32, _I32, _I32, _D,
                                      _D, _D, _U32]
        lib.bm_level_noalloc.argtypes = [vp, ctypes.c_int32] + [vp] * 7 + [ctypes.c_int32] + [vp] * 6
        lib.bm_set_launch.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        lib.bm_set_colour_lds.argtypes = [vp, ctypes.c_int32]
        lib.bm_set_wide.argtypes = [vp, ctypes.c_int32]
        lib.bm_set_device_only.argtypes = [vp, ctypes.c_int32]
        lib.bm_lds_limit.argtypes = [vp]
        lib.bm_lds_limit.restype = ctypes.c_int64
        world_signatures(lib)
        _lib = lib
    return _lib


class Bodies(ctypes.Structure):
    """bm_bodies: addBody's resolved columns, one entry a body."""
    _fields_ = [("n", ctypes.c_int32), ("type", _U8), ("shape", _U8), ("lock", _U8)] + [
        (f, _D) for f in ("position", "rotation", "velocity", "angular", "dims", "inertia", "inv_inertia", "mass",
                          "inv_mass", "gravity_scale", "damping", "angular_damping", "size", "reach", "friction",
                          "restitution")] + [("hull", _I32), ("framed", _U8), ("offset", _D), ("frame", _D),
                                                         ("field", _I32)]


class WorldState(ctypes.Structure):
    _fields_ = [("top", ctypes.c_int32), ("awake", ctypes.c_int32), ("contacts", ctypes.c_int32),
                ("pairs", ctypes.c_int32), ("frozen", ctypes.c_int32), ("impacts", ctypes.c_int64),
                ("catches", ctypes.c_int64),
                ("transforms", ctypes.POINTER(ctypes.c_float)), ("position", _D), ("rotation", _D), ("velocity", _D),
                ("angular", _D), ("rows", ctypes.c_int64), ("soft_bytes", ctypes.c_double), ("soft_ms", ctypes.c_double),
                ("previous", ctypes.POINTER(ctypes.c_float))]


class WorldEvents(ctypes.Structure):
    """bm_world_events: the last step's contact events, as the world holds them."""
    _fields_ = [("count", ctypes.c_int32), ("type", _U8), ("a", _I32), ("b", _I32), ("point", _D), ("normal", _D),
                ("speed", _D)]


def world_signatures(lib):
    """Sets the bm_world_* signatures on a library that exports them."""
    vp = ctypes.c_void_p
    lib.bm_world_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, _D, ctypes.c_int32, ctypes.c_int32]
    lib.bm_world_destroy.argtypes = [vp]
    lib.bm_world_error.argtypes = [vp]
    lib.bm_world_error.restype = ctypes.c_char_p
    lib.bm_world_add.argtypes = [vp, ctypes.POINTER(Bodies)]
    lib.bm_world_step.argtypes = [vp, ctypes.c_double]
    lib.bm_world_read.argtypes = [vp, ctypes.POINTER(WorldState)]
    lib.bm_world_moved.argtypes = [vp, ctypes.c_int32, _I32, _I32]
    lib.bm_world_tap.argtypes = [vp, COLUMNS_TAP, vp]
    lib.bm_world_events_of.argtypes = [vp, ctypes.POINTER(WorldEvents)]
    lib.bm_world_set_broadphase.argtypes = [vp, ctypes.c_int32]
    lib.bm_world_set_parallel.argtypes = [vp, ctypes.c_int32]
    lib.bm_world_set_hulls.argtypes = [vp, ctypes.c_int32, _I32, _I32, _I32, _D, _I32, _I32, _D]
    lib.bm_world_set_fields.argtypes = [vp, ctypes.c_int32, _I32, _I32, _D, ctypes.POINTER(ctypes.c_int64), _D]
    lib.bm_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.c_int32]
    lib.bm_batch_destroy.argtypes = [vp]
    lib.bm_batch_error.argtypes = [vp]
    lib.bm_batch_error.restype = ctypes.c_char_p
    lib.bm_batch_world.argtypes = [vp, ctypes.POINTER(vp), _D, ctypes.c_int32, ctypes.c_int32]
    lib.bm_batch_step.argtypes = [vp, ctypes.c_double]
    lib.bm_batch_set_graph.argtypes = [vp, ctypes.c_int32]
    lib.bm_batch_set_launch.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    lib.bm_batch_graph_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.bm_world_times.argtypes = [vp, _D]
    lib.bm_batch_times.argtypes = [vp, _D]
    lib.bm_world_record.argtypes = [vp, ctypes.c_int32]
    lib.bm_world_recorded.argtypes = [vp]
    lib.bm_adjoint_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.POINTER(vp), ctypes.c_int32,
                                      ctypes.c_int32]
    lib.bm_adjoint_destroy.argtypes = [vp]
    lib.bm_adjoint_error.argtypes = [vp]
    lib.bm_adjoint_error.restype = ctypes.c_char_p
    lib.bm_adjoint_size.argtypes = [vp, _I32, _I32]
    lib.bm_adjoint_recorded.argtypes = [vp, _D, _D, _D, _D]
    lib.bm_adjoint_forward.argtypes = [vp, _D, _D, _D, _D, _D]
    lib.bm_adjoint_backward.argtypes = [vp, _D, _D, _D, _D, _D]
    lib.bm_adjoint_times.argtypes = [vp, _D, _D]
    lib.bm_adjoint_branches.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    lib.bm_adjoint_held.argtypes = [vp, ctypes.POINTER(Held)]


class Held(ctypes.Structure):
    """bm_held: a held trajectory as the device takes it."""
    _fields_ = ([(f, ctypes.c_int32) for f in ("worlds", "steps", "every", "bodies", "rows", "first")]
                + [(f, _I32) for f in ("body_offset", "row_offset", "starts", "narrow", "substeps")]
                + [(f, _D) for f in ("dt", "gravity")] + [(f, _U8) for f in ("type", "shape", "tumbles", "capped")]
                + [(f, _D) for f in ("rounded", "segment", "gravity_scale", "damping", "angular_damping", "mass0",
                                     "inv_mass0", "inertia0", "inv_inertia0", "left", "ceiling")]
                + [(f, _I32) for f in ("a", "b", "src")] + [("count", _U8), ("pid", ctypes.POINTER(ctypes.c_uint32))]
                + [(f, _D) for f in ("local_normal", "ends", "shift", "tolerance", "patch", "store_impulse",
                                     "store_friction", "store_twist", "store_slip")])


class BearingError(RuntimeError):
    pass


class Context:
    """One bm_context on a HIP device (or on `lib`, another library of the same calls)."""

    def __init__(self, device=0, lib=None):
        if lib is None:
            lib = load()
        else:
            context_signatures(lib)
        self.lib = lib
        self.ptr = ctypes.c_void_p()
        rc = self.lib.bm_create(ctypes.byref(self.ptr), device)
        if rc != 0:
            raise BearingError(f"bm_create: no HIP device {device} (status {rc})")

    def check(self, rc, what):
        if rc != 0:
            raise BearingError(f"{what}: status {rc}: {self.lib.bm_last_error(self.ptr).decode()}")

    def set_launch(self, shape, resident=True):
        self.check(self.lib.bm_set_launch(self.ptr, shape, 1 if resident else 0), "bm_set_launch")

    def set_colour_lds(self, allow=True):
        """bm_set_colour_lds: True (1) colour in LDS where a world fits, False (0) the HBM path,
        "stream" (2) a single world's manifolds streamed through one wave and radix-sorted."""
        mode = 2 if allow == "stream" else 1 if allow else 0
        self.check(self.lib.bm_set_colour_lds(self.ptr, mode), "bm_set_colour_lds")

    def set_wide(self, min_manifolds):
        """bm_set_wide: a single world of at least min_manifolds manifolds takes the grid-wide
        soft step (0: never)."""
        self.check(self.lib.bm_set_wide(self.ptr, min_manifolds), "bm_set_wide")

    def lds_limit(self):
        return self.lib.bm_lds_limit(self.ptr)

    def close(self):
        if self.ptr:
            self.lib.bm_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_hulls(ctx, table):
    """bm_set_hulls with a hull table: a dict of hullV, hullF, hullL (offsets, n + 1 each),
    vertices, start, loops, planes."""
    i32 = lambda x: np.ascontiguousarray(x, np.int32)
    f64 = lambda x: np.ascontiguousarray(x, np.float64)
    hv, hf, hl = i32(table["hullV"]), i32(table["hullF"]), i32(table["hullL"])
    v, st, lp, pl = f64(table["vertices"]), i32(table["start"]), i32(table["loops"]), f64(table["planes"])
    ctx.check(ctx.lib.bm_set_hulls(ctx.ptr, len(hv) - 1, hv.ctypes.data_as(_I32), hf.ctypes.data_as(_I32),
                                   hl.ctypes.data_as(_I32), v.ctypes.data_as(_D), st.ctypes.data_as(_I32),
                                   lp.ctypes.data_as(_I32), pl.ctypes.data_as(_D)), "bm_set_hulls")


def set_fields(ctx, columns, rows, spacing, heights):
    """bm_set_fields: heightfields by number, field f columns[f] x rows[f] samples spacing[f]
    apart, heights[f] its samples row by row."""
    n = len(columns)
    cols, rws = np.ascontiguousarray(columns, np.int32), np.ascontiguousarray(rows, np.int32)
    sp = np.ascontiguousarray(spacing, np.float64)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(h) for h in heights])
    flat = np.ascontiguousarray(np.concatenate([np.asarray(h, np.float64) for h in heights]) if n else np.zeros(1))
    ctx.check(ctx.lib.bm_set_fields(ctx.ptr, n, cols.ctypes.data_as(_I32), rws.ctypes.data_as(_I32),
                                    sp.ctypes.data_as(_D), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    flat.ctypes.data_as(_D)), "bm_set_fields")


def level(ctx, field, position, axes, shapes, boxes, margin, tolerance, cap):
    """bm_level_host (touchLevel) for each shape record against its field: manifolds a
    shape, then count, normal, points, separations, ids in cap slots a shape."""
    n = len(margin)
    f64 = lambda x: np.ascontiguousarray(x, np.float64)
    field = np.ascontiguousarray(field, np.int32)
    position, axes, shapes, boxes, margin, tolerance = map(f64, (position, axes, shapes, boxes, margin, tolerance))
    manifolds = np.zeros(n, np.int32)
    count = np.zeros(cap * n, np.int32)
    normal = np.zeros(3 * cap * n)
    points = np.zeros(12 * cap * n)
    seps = np.zeros(4 * cap * n)
    ids = np.zeros(4 * cap * n, np.uint32)
    p = lambda x, t: x.ctypes.data_as(t)
    ctx.check(ctx.lib.bm_level_host(ctx.ptr, n, p(field, _I32), p(position, _D), p(axes, _D), p(shapes, _D),
                                    p(boxes, _D), p(margin, _D), p(tolerance, _D), cap, p(manifolds, _I32),
                                    p(count, _I32), p(normal, _D), p(points, _D), p(seps, _D), p(ids, _U32)),
              "bm_level_host")
    return manifolds, count, normal, points, seps, ids


def level_noalloc(ctx, n, cap, field, position, axes, shapes, boxes, margin, tolerance, manifolds, count, normal,
                  points, separations, ids):
    """bm_level_noalloc: as level(), every argument after cap a device address (an int: the
    caller's device memory, in level()'s layout); the outputs are written there."""
    vp = ctypes.c_void_p
    ctx.check(ctx.lib.bm_level_noalloc(ctx.ptr, n, *[vp(a) for a in (field, position, axes, shapes, boxes, margin)],
                                       vp(tolerance), cap,
                                       *[vp(a) for a in (manifolds, count, normal, points, separations, ids)]),
              "bm_level_noalloc")


def collide(ctx, a, b, margin, tolerance, with_loss=False):
    """bm_collide_host on flat shape records: (count, normal, points, separations, ids) per
    pair, and each manifold's loss after them when asked."""
    n = len(margin)
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    margin = np.ascontiguousarray(margin, np.float64)
    tolerance = np.ascontiguousarray(tolerance, np.float64)
    count = np.zeros(n, np.int32)
    normal = np.zeros(3 * n)
    points = np.zeros(12 * n)
    separations = np.zeros(4 * n)
    ids = np.zeros(4 * n, np.uint32)
    loss = np.zeros(n)
    p = lambda x, k: x.ctypes.data_as(k)
    ctx.check(ctx.lib.bm_collide_host(ctx.ptr, n, p(a, _D), p(b, _D), p(margin, _D), p(tolerance, _D), p(count, _I32),
                                      p(normal, _D), p(points, _D), p(separations, _D), p(ids, _U32), p(loss, _D)),
              "bm_collide_host")
    if with_loss:
        return count, normal, points, separations, ids, loss
    return count, normal, points, separations, ids


def meet(ctx, a, b, pair):
    """bm_meet_host on body records (32 numbers) and pair records (margin, tolerance, dt,
    gravity 3): (count, normal, points, separations, ids, shift, flags) per pair."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    pair = np.ascontiguousarray(pair, np.float64)
    n = pair.size // 6
    count, flags = np.zeros(n, np.int32), np.zeros(n, np.int32)
    normal, points, separations = np.zeros(3 * n), np.zeros(12 * n), np.zeros(4 * n)
    ids, shift, loss = np.zeros(4 * n, np.uint32), np.zeros(12 * n), np.zeros(n)
    ctx.check(ctx.lib.bm_meet_host(ctx.ptr, n, a.ctypes.data_as(_D), b.ctypes.data_as(_D), pair.ctypes.data_as(_D),
                                   count.ctypes.data_as(_I32), normal.ctypes.data_as(_D), points.ctypes.data_as(_D),
                                   separations.ctypes.data_as(_D), ids.ctypes.data_as(_U32), shift.ctypes.data_as(_D),
                                   flags.ctypes.data_as(_I32), loss.ctypes.data_as(_D)), "bm_meet_host")
    return count, normal, points, separations, ids, shift, flags, loss


def meet_packed(ctx, table, ia, ib, pair):
    """bm_meet_packed on a table of body records (32 numbers each), each pair two row numbers
    into it, and pair records: (count, flags, row, normal, points, separations, ids, shift,
    loss), the last six packed by row (the pairs with count > 0, in pair order)."""
    table = np.ascontiguousarray(table, np.float64)
    ia = np.ascontiguousarray(ia, np.int32)
    ib = np.ascontiguousarray(ib, np.int32)
    pair = np.ascontiguousarray(pair, np.float64)
    n, bodies = pair.size // 6, table.size // 32
    count, flags, row = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
    normal, points, separations = np.zeros(3 * n), np.zeros(12 * n), np.zeros(4 * n)
    ids, shift, loss = np.zeros(4 * n, np.uint32), np.zeros(12 * n), np.zeros(n)
    ctx.check(ctx.lib.bm_meet_packed(ctx.ptr, bodies, table.ctypes.data_as(_D), n, ia.ctypes.data_as(_I32),
                                     ib.ctypes.data_as(_I32), pair.ctypes.data_as(_D), count.ctypes.data_as(_I32),
                                     flags.ctypes.data_as(_I32), row.ctypes.data_as(_I32), normal.ctypes.data_as(_D),
                                     points.ctypes.data_as(_D), separations.ctypes.data_as(_D),
                                     ids.ctypes.data_as(_U32), shift.ctypes.data_as(_D), loss.ctypes.data_as(_D)),
              "bm_meet_packed")
    found = int((count > 0).sum())
    return (count, flags, row, normal[:3 * found], points[:12 * found], separations[:4 * found], ids[:4 * found],
            shift[:12 * found], loss[:found])


class Batch:
    """Worlds' soft-step columns, concatenated. Each world is a dict by bm_columns field
    name (per-body, per-manifold, per-pair arrays plus 'awake_list', 'starts', 'solved',
    'substeps', 'dt', 'gravity')."""

    def __init__(self, worlds):
        self.worlds = worlds
        n = len(worlds)
        counts = {u: [] for u in PER_WORLD_UNITS}
        for w in worlds:
            counts["B"].append(len(w["type"]))
            counts["A"].append(len(w["awake_list"]))
            counts["M"].append(len(w["points"]))
            counts["P"].append(len(w["pair_points"]))
        self.offsets = {u: np.concatenate([[0], np.cumsum(c)]).astype(np.int32) for u, c in counts.items()}
        self.arrays = {}
        for f, kind, unit, per in FIELDS:
            dt = _NP[kind]
            if unit == "W1":
                key = {"body_offset": "B", "manifold_offset": "M", "pair_offset": "P", "awake_offset": "A"}[f]
                arr = self.offsets[key]
            elif unit in ("W", "W3", "S"):
                arr = np.concatenate([np.atleast_1d(np.asarray(w[f], dtype=dt)) for w in worlds]).astype(dt)
            else:
                arr = np.concatenate([np.asarray(w[f], dtype=dt).reshape(-1) for w in worlds]).astype(dt)
                expect = int(self.offsets[unit][-1]) * per
                if arr.size != expect:
                    raise ValueError(f"Batch: {f} has {arr.size} numbers, expected {expect}")
            self.arrays[f] = np.ascontiguousarray(arr)
        self.n = n
        self.struct = Columns()
        self.struct.n_worlds = n
        for f, kind, _, _ in FIELDS:
            setattr(self.struct, f, self.arrays[f].ctypes.data_as(kind))

    def world(self, i):
        """World i's columns after the step (views into the batch arrays)."""
        out = {}
        for f, kind, unit, per in FIELDS:
            if unit in PER_WORLD_UNITS:
                o = self.offsets[unit]
                out[f] = self.arrays[f][o[i] * per:o[i + 1] * per]
        return out

    def upload(self, ctx):
        ctx.check(ctx.lib.bm_upload(ctx.ptr, ctypes.byref(self.struct)), "bm_upload")

    def download(self, ctx):
        ctx.check(ctx.lib.bm_download(ctx.ptr, ctypes.byref(self.struct)), "bm_download")

    def step(self, ctx):
        """bm_soft_step on the uploaded batch; returns the kernel's milliseconds."""
        ctx.check(ctx.lib.bm_soft_step(ctx.ptr), "bm_soft_step")
        return ctx.lib.bm_last_kernel_ms(ctx.ptr)

    def colour(self, ctx):
        """bm_colour on the uploaded batch (manifolds in the order narrow() found them)."""
        ctx.check(ctx.lib.bm_colour(ctx.ptr), "bm_colour")

    def step_host(self, ctx):
        ctx.check(ctx.lib.bm_soft_step_host(ctx.ptr, ctypes.byref(self.struct)), "bm_soft_step_host")


# addBody's resolved columns (bm_bodies field, numbers a body, dtype).
BODY_COLUMNS = [("type", 1, np.uint8), ("shape", 1, np.uint8), ("lock", 1, np.uint8), ("position", 3, np.float64),
                ("rotation", 4, np.float64), ("velocity", 3, np.float64), ("angular", 3, np.float64),
                ("dims", 3, np.float64), ("inertia", 3, np.float64), ("inv_inertia", 3, np.float64),
                ("mass", 1, np.float64), ("inv_mass", 1, np.float64), ("gravity_scale", 1, np.float64),
                ("damping", 1, np.float64), ("angular_damping", 1, np.float64), ("size", 1, np.float64),
                ("reach", 1, np.float64), ("friction", 1, np.float64), ("restitution", 1, np.float64),
                ("hull", 1, np.int32), ("framed", 1, np.uint8), ("offset", 3, np.float64), ("frame", 4, np.float64),
                ("field", 1, np.int32)]
# Columns a world of unframed boxes, spheres and capsules can leave out (passed as NULL).
OPTIONAL_COLUMNS = {"hull", "framed", "offset", "frame", "field"}


class World:
    """A whole world stepped as the reference's World.step steps it (bm_world_*).

    `lib` is the library to bind, the built one by default."""

    def __init__(self, device=0, gravity=(0.0, -9.81, 0.0), substeps=8, sleeping=True, lib=None, batch=None):
        if batch is not None:
            lib = batch.lib
        elif lib is None:
            lib = load()
        else:
            world_signatures(lib)
        self.lib = lib
        self.ptr = ctypes.c_void_p()
        self.batch = batch
        g = np.ascontiguousarray(gravity, np.float64)
        if batch is not None:
            batch.check(lib.bm_batch_world(batch.ptr, ctypes.byref(self.ptr), g.ctypes.data_as(_D), substeps,
                                           1 if sleeping else 0), "bm_batch_world")
            return
        rc = lib.bm_world_create(ctypes.byref(self.ptr), device, g.ctypes.data_as(_D), substeps, 1 if sleeping else 0)
        if rc != 0:
            raise BearingError(f"bm_world_create: no HIP device {device} (status {rc})")

    def check(self, rc, what):
        if rc != 0:
            raise BearingError(f"{what}: status {rc}: {self.lib.bm_world_error(self.ptr).decode()}")

    BROADPHASE = {"host": 0, "device": 1, "auto": 2}

    def set_parallel(self, from_):
        """bm_world_set_parallel: from how many pairs (bodies, rows) a pass's work goes to
        the world's host threads."""
        self.check(self.lib.bm_world_set_parallel(self.ptr, int(from_)), "bm_world_set_parallel")

    def set_broadphase(self, mode):
        """Where the broadphase's passes run: "host", "device" or "auto" (bm_world_set_broadphase)."""
        self.check(self.lib.bm_world_set_broadphase(self.ptr, self.BROADPHASE[mode]), "bm_world_set_broadphase")

    def record(self, on=True):
        """bm_world_record: keep the steps from now on for an Adjoint (sleeping must be off)."""
        self.check(self.lib.bm_world_record(self.ptr, 1 if on else 0), "bm_world_record")

    def recorded(self):
        return int(self.lib.bm_world_recorded(self.ptr))

    def add(self, columns):
        """addBody for each body of `columns` (name -> array, names as BODY_COLUMNS)."""
        n = len(columns["type"])
        keep, b = [], Bodies(n=n)
        for name, per, dtype in BODY_COLUMNS:
            if name in OPTIONAL_COLUMNS and columns.get(name) is None:
                continue
            arr = np.ascontiguousarray(columns[name], dtype)
            if arr.size != per * n:
                raise ValueError(f"{name}: {arr.size} numbers for {n} bodies")
            keep.append(arr)
            setattr(b, name, arr.ctypes.data_as(_U8 if dtype == np.uint8 else _I32 if dtype == np.int32 else _D))
        self.check(self.lib.bm_world_add(self.ptr, ctypes.byref(b)), "bm_world_add")

    def set_hulls(self, table):
        """bm_world_set_hulls with a hull table: a dict of hullV, hullF, hullL (offsets, n + 1
        each), vertices, start, loops, planes, as set_hulls takes."""
        i32 = lambda x: np.ascontiguousarray(x, np.int32)
        f64 = lambda x: np.ascontiguousarray(x, np.float64)
        hv, hf, hl = i32(table["hullV"]), i32(table["hullF"]), i32(table["hullL"])
        v, st, lp, pl = f64(table["vertices"]), i32(table["start"]), i32(table["loops"]), f64(table["planes"])
        self.check(self.lib.bm_world_set_hulls(self.ptr, len(hv) - 1, hv.ctypes.data_as(_I32), hf.ctypes.data_as(_I32),
                                               hl.ctypes.data_as(_I32), v.ctypes.data_as(_D), st.ctypes.data_as(_I32),
                                               lp.ctypes.data_as(_I32), pl.ctypes.data_as(_D)), "bm_world_set_hulls")

    def set_fields(self, columns, rows, spacing, heights):
        """bm_world_set_fields: field f columns[f] x rows[f] samples spacing[f] apart, heights[f]
        its samples row by row."""
        n = len(columns)
        cols, rws = np.ascontiguousarray(columns, np.int32), np.ascontiguousarray(rows, np.int32)
        sp = np.ascontiguousarray(spacing, np.float64)
        off = np.zeros(n + 1, np.int64)
        off[1:] = np.cumsum([len(h) for h in heights])
        flat = np.ascontiguousarray(np.concatenate([np.asarray(h, np.float64) for h in heights]) if n else np.zeros(1))
        self.check(self.lib.bm_world_set_fields(self.ptr, n, cols.ctypes.data_as(_I32), rws.ctypes.data_as(_I32),
                                                sp.ctypes.data_as(_D), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                flat.ctypes.data_as(_D)), "bm_world_set_fields")

    def step(self, dt):
        self.check(self.lib.bm_world_step(self.ptr, dt), "bm_world_step")

    def state(self):
        s = WorldState()
        self.check(self.lib.bm_world_read(self.ptr, ctypes.byref(s)), "bm_world_read")
        return s

    def read(self):
        """The counts, and transforms (float32, 7 a slot), position, rotation, velocity, angular."""
        top = self.state().top
        out = {"transforms": np.zeros(7 * top, np.float32), "previous": np.zeros(7 * top, np.float32),
               "position": np.zeros(3 * top), "rotation": np.zeros(4 * top),
               "velocity": np.zeros(3 * top), "angular": np.zeros(3 * top)}
        s = WorldState()
        s.transforms = out["transforms"].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        s.previous = out["previous"].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        for name in ("position", "rotation", "velocity", "angular"):
            setattr(s, name, out[name].ctypes.data_as(_D))
        self.check(self.lib.bm_world_read(self.ptr, ctypes.byref(s)), "bm_world_read")
        for name in ("top", "awake", "contacts", "pairs", "frozen", "impacts", "catches", "rows", "soft_bytes",
                     "soft_ms"):
            out[name] = getattr(s, name)
        return out

    def moved(self, since):
        """world.moved(since): the slots that moved after step `since`, in the reference's order."""
        out = np.zeros(max(self.state().top, 1), np.int32)
        n = np.zeros(1, np.int32)
        self.check(self.lib.bm_world_moved(self.ptr, since, out.ctypes.data_as(_I32), n.ctypes.data_as(_I32)),
                   "bm_world_moved")
        return out[:int(n[0])].copy()

    def events(self):
        """The last step's contact events (world.contacts(), src/world.js:690-723) as an (n, 10)
        float64 array, a row [type, slot a, slot b, point 3, normal 3, speed] an event (type 0
        begin, 1 end; an end's point, normal and speed 0), in the order the step wrote them."""
        e = WorldEvents()
        self.check(self.lib.bm_world_events_of(self.ptr, ctypes.byref(e)), "bm_world_events_of")
        n = e.count
        out = np.zeros((n, 10))
        if n:
            out[:, 0] = np.ctypeslib.as_array(e.type, (n,))
            out[:, 1] = np.ctypeslib.as_array(e.a, (n,))
            out[:, 2] = np.ctypeslib.as_array(e.b, (n,))
            out[:, 3:6] = np.ctypeslib.as_array(e.point, (3 * n,)).reshape(n, 3)
            out[:, 6:9] = np.ctypeslib.as_array(e.normal, (3 * n,)).reshape(n, 3)
            out[:, 9] = np.ctypeslib.as_array(e.speed, (n,))
        return out

    def capture_next(self):
        """Arms bm_world_tap for the next step alone: after it, `captured` holds the columns that
        step handed the device (copy_columns), before the device changed them."""
        self.captured = None

        def tap(_user, cols):
            self.captured = copy_columns(cols.contents)
            self.lib.bm_world_tap(self.ptr, COLUMNS_TAP(), None)

        self._tap = COLUMNS_TAP(tap)  # kept alive while the library holds it
        self.check(self.lib.bm_world_tap(self.ptr, self._tap, None), "bm_world_tap")

    TIMES = ["prepare", "narrowphase_device", "narrow_host", "freeze", "contact_device", "catch_sleep"]

    def times(self):
        """Milliseconds the steps so far spent in each part (bm_world_times), by name."""
        out = np.zeros(len(self.TIMES))
        self.check(self.lib.bm_world_times(self.ptr, out.ctypes.data_as(_D)), "bm_world_times")
        return dict(zip(self.TIMES, out.tolist()))

    def close(self):
        if self.ptr and self.batch is None:
            self.lib.bm_world_destroy(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WorldBatch:
    """Worlds stepped together (bm_batch_*): each as World steps it, the device shared by
    all of them. world() makes a World of the batch; step(dt) steps every one."""

    def __init__(self, device=0, threads=0, lib=None):
        if lib is None:
            lib = load()
        else:
            world_signatures(lib)
        self.lib = lib
        self.ptr = ctypes.c_void_p()
        self.worlds = []
        rc = lib.bm_batch_create(ctypes.byref(self.ptr), device, threads)
        if rc != 0:
            raise BearingError(f"bm_batch_create: no HIP device {device} (status {rc})")

    def check(self, rc, what):
        if rc != 0:
            raise BearingError(f"{what}: status {rc}: {self.lib.bm_batch_error(self.ptr).decode()}")

    def world(self, gravity=(0.0, -9.81, 0.0), substeps=8, sleeping=True):
        w = World(gravity=gravity, substeps=substeps, sleeping=sleeping, batch=self)
        self.worlds.append(w)
        return w

    def step(self, dt):
        self.check(self.lib.bm_batch_step(self.ptr, dt), "bm_batch_step")

    def times(self):
        """Milliseconds the batch's steps spent in each part (bm_batch_times), by name."""
        out = np.zeros(len(World.TIMES))
        self.check(self.lib.bm_batch_times(self.ptr, out.ctypes.data_as(_D)), "bm_batch_times")
        return dict(zip(World.TIMES, out.tolist()))

    def set_graph(self, on):
        """Whether the contact phase goes to the device as a HIP graph (bm_batch_set_graph)."""
        self.check(self.lib.bm_batch_set_graph(self.ptr, 1 if on else 0), "bm_batch_set_graph")

    def set_launch(self, shape, resident=True):
        """The soft step's launch shape (0, 1, 2) and velocity residency (bm_batch_set_launch)."""
        self.check(self.lib.bm_batch_set_launch(self.ptr, int(shape), 1 if resident else 0), "bm_batch_set_launch")

    def graph_counts(self):
        """{launches, made, updated}: the contact phase's graphs so far (bm_batch_graph_counts)."""
        out = np.zeros(3, np.int64)
        self.check(self.lib.bm_batch_graph_counts(self.ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
                   "bm_batch_graph_counts")
        return dict(zip(("launches", "made", "updated"), out.tolist()))

    def close(self):
        if self.ptr:
            for w in self.worlds:
                w.ptr = ctypes.c_void_p()
            self.lib.bm_batch_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HELD_STATE = 13  # p 3, q 4, v 3, w 3 a body (BM_HELD_STATE)


class Adjoint:
    """The held trajectory of recorded worlds and its gradients (bm_adjoint_*): forward()
    steps it from an initial state with given masses, frictions and restitutions;
    backward() gives the adjoints of those from the final state's.

    Arrays are by body slot, the worlds' slots one after another."""

    def __init__(self, worlds, every=10, device=0, lib=None):
        self.lib = lib if lib is not None else load()
        if lib is not None:
            world_signatures(lib)
        self.ptr = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(worlds))(*[w.ptr for w in worlds])
        self.worlds = len(worlds)
        rc = self.lib.bm_adjoint_create(ctypes.byref(self.ptr), device, arr, len(worlds), every)
        if rc != 0:
            msg = self.lib.bm_adjoint_error(self.ptr).decode() if self.ptr else ""
            self.lib.bm_adjoint_destroy(self.ptr)
            self.ptr = None
            raise BearingError(f"bm_adjoint_create: status {rc}: {msg}")
        b, t = np.zeros(1, np.int32), np.zeros(1, np.int32)
        self.lib.bm_adjoint_size(self.ptr, b.ctypes.data_as(_I32), t.ctypes.data_as(_I32))
        self.bodies, self.steps = int(b[0]), int(t[0])
        # Forward passes run so far: backward() walks the last one's checkpoints, so a caller
        # holding an earlier pass's count knows to run that pass again first.
        self.forwards = 0

    def check(self, rc, what):
        if rc != 0:
            raise BearingError(f"{what}: status {rc}: {self.lib.bm_adjoint_error(self.ptr).decode()}")

    def recorded(self):
        """The recorded initial state (bodies x 13) and each slot's mass, friction, restitution."""
        n = self.bodies
        out = {"state": np.zeros((n, HELD_STATE)), "mass": np.zeros(n), "friction": np.zeros(n),
               "restitution": np.zeros(n)}
        self.check(self.lib.bm_adjoint_recorded(self.ptr, *[out[k].ctypes.data_as(_D) for k in
                                                          ("state", "mass", "friction", "restitution")]),
                   "bm_adjoint_recorded")
        return out

    def forward(self, state, mass, friction, restitution):
        n = self.bodies
        ins = [np.ascontiguousarray(a, np.float64) for a in (state, mass, friction, restitution)]
        if ins[0].size != n * HELD_STATE or any(a.size != n for a in ins[1:]):
            raise ValueError("forward: arrays of the wrong size")
        out = np.zeros((n, HELD_STATE))
        self.forwards += 1
        self.check(self.lib.bm_adjoint_forward(self.ptr, *[a.ctypes.data_as(_D) for a in ins],
                                               out.ctypes.data_as(_D)), "bm_adjoint_forward")
        return out

    def backward(self, state_bar):
        n = self.bodies
        sb = np.ascontiguousarray(state_bar, np.float64)
        if sb.size != n * HELD_STATE:
            raise ValueError("backward: state_bar of the wrong size")
        out = {"state": np.zeros((n, HELD_STATE)), "mass": np.zeros(n), "friction": np.zeros(n),
               "restitution": np.zeros(n)}
        self.check(self.lib.bm_adjoint_backward(self.ptr, sb.ctypes.data_as(_D),
                                                *[out[k].ctypes.data_as(_D) for k in
                                                  ("state", "mass", "friction", "restitution")]),
                   "bm_adjoint_backward")
        return out

    def held(self):
        """bm_adjoint_held: the held trajectory's description, every array copied out."""
        h = Held()
        self.check(self.lib.bm_adjoint_held(self.ptr, ctypes.byref(h)), "bm_adjoint_held")
        W, T, B, R, F = h.worlds, h.steps, h.bodies, h.rows, h.first
        sizes = {"body_offset": W + 1, "row_offset": T * W + 1, "starts": T * W * 35, "narrow": R, "substeps": W,
                 "dt": W, "gravity": 3 * W, "type": B, "shape": B, "tumbles": B, "capped": B, "rounded": B,
                 "segment": B, "gravity_scale": B, "damping": B, "angular_damping": B, "mass0": B, "inv_mass0": B,
                 "inertia0": 3 * B, "inv_inertia0": 3 * B, "left": 3 * B, "ceiling": B, "a": R, "b": R, "src": R,
                 "count": R, "pid": 4 * R, "local_normal": 3 * R, "ends": 24 * R, "shift": 12 * R, "tolerance": R,
                 "patch": R, "store_impulse": 4 * F, "store_friction": 3 * F, "store_twist": 3 * F,
                 "store_slip": 3 * F}
        out = {k: getattr(h, k) for k in ("worlds", "steps", "every", "bodies", "rows", "first")}
        for name, n in sizes.items():
            ptr = getattr(h, name)
            out[name] = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0)
        return out

    def branches(self):
        """bm_adjoint_branches: each world's record of the branches the last forward pass took."""
        n = int(self.worlds)
        out = np.zeros(n, np.uint64)
        self.check(self.lib.bm_adjoint_branches(self.ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))),
                   "bm_adjoint_branches")
        return out

    def times(self):
        f, b = np.zeros(1), np.zeros(1)
        self.check(self.lib.bm_adjoint_times(self.ptr, f.ctypes.data_as(_D), b.ctypes.data_as(_D)),
                   "bm_adjoint_times")
        return float(f[0]), float(b[0])

    def close(self):
        if self.ptr:
            self.lib.bm_adjoint_destroy(self.ptr)
            self.ptr = None
