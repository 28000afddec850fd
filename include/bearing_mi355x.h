/*
 * bearing_mi355x — the contact pipeline of the Bearing physics engine on one AMD
 * Instinct MI355X (gfx950). C ABI: plain pointers and sizes, int status codes,
 * no C++ exceptions or torch types across it.
 *
 * The reference (bearing-physics, JavaScript) keeps a world as typed-array
 * columns by body slot (src/world.js:228-311), a step's manifolds as columns by
 * manifold and by point (src/contact.js:149-258, four point slots a manifold)
 * and the pair list's remembered impulses by pair (src/pairs.js:40-110). This
 * interface takes exactly those columns, in exactly that layout, so an N-API
 * binding hands the reference's own typed arrays through unchanged.
 *
 * A batch holds one or more independent worlds. Column arrays of a batch are
 * the worlds' columns one after another; the offset tables say where each
 * world's bodies, manifolds, pairs and awake list begin. Slots, manifold
 * numbers and pair numbers inside the columns (a, b, pair, row, awake_list)
 * are each world's own, as the reference numbers them.
 *
 * Threading: a context is used by one thread at a time; two contexts are
 * independent. Host pointers are read or written only during the call that
 * takes them and never kept.
 */
#ifndef BEARING_MI355X_H
#define BEARING_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BM_OK 0
#define BM_ERR_ARGUMENT 1   /* a size, pointer or offset table is invalid */
#define BM_ERR_DEVICE 2     /* a HIP call failed; bm_last_error says which */
#define BM_ERR_UNSUPPORTED 3 /* a manifold of a mesh or compound pair (sub >= 0) */

#define BM_COLOURS 32            /* src/graph.js:67 */
#define BM_STARTS (BM_COLOURS + 3) /* colour runs, overflow, resting, end: src/contact.js:160 */

/* The columns of a batch of worlds, host or device side. Arrays marked [in]
 * are read, [out] written, [io] both. Per-unit widths follow the reference. */
typedef struct bm_columns {
  int32_t n_worlds;
  /* offset tables, n_worlds + 1 entries each */
  const int32_t* body_offset;
  const int32_t* manifold_offset;
  const int32_t* pair_offset;
  const int32_t* awake_offset;
  /* per world */
  int32_t* starts;           /* [io] BM_STARTS per world: contacts.starts after colour() */
  int32_t* solved;           /* [io] contacts.solved: rows before the resting run */
  const int32_t* substeps;   /* [in] world.substeps */
  const double* dt;          /* [in] the step's dt */
  const double* gravity;     /* [in] 3 per world */
  /* bodies, by slot (src/world.js COLUMNS) */
  double* position;          /* [io] 3 */
  double* rotation;          /* [io] 4 */
  double* velocity;          /* [io] 3 */
  double* angular;           /* [io] 3 */
  double* start;             /* [out] 3: where each awake body began the step */
  double* start_q;           /* [out] 4 */
  double* inv_i;             /* [io] 6: world inverse inertia, written for awake bodies */
  double* tumbled;           /* [io] 3 */
  double* ceiling;           /* [io] 1 */
  const double* dims;        /* [in] 3 */
  const double* inv_mass;    /* [in] 1 */
  const double* inertia;     /* [in] 3 */
  const double* inv_inertia; /* [in] 3 */
  const double* axes;        /* [in] 9: the rotation's columns as _bounds left them */
  const double* gravity_scale, *damping, *angular_damping; /* [in] 1 each */
  const uint8_t* type, *shape, *frozen, *lock, *framed, *awake; /* [in] 1 each */
  const int32_t* joint_head; /* [in] 1 */
  const int32_t* awake_list; /* [in] awake_offset-indexed */
  /* manifolds, by row as colour() laid them out (src/graph.js arrange); before
   * bm_colour, in the order narrow() found them */
  int32_t* a, *b, *pair, *sub; /* [io] 1 (moved to their rows by bm_colour) */
  int32_t* row;              /* [io] 1: the row each found manifold went to */
  int32_t* colour;           /* [io] 1: colour of the manifold in each row (32 overflow, 33 resting) */
  int32_t* order;            /* [io] 1: which found manifold each row holds */
  uint8_t* points;           /* [io] 1 */
  uint8_t* sliding;          /* [out] 1 */
  double* normal, *t1, *t2;  /* [io] 3 */
  double* friction, *restitution, *tolerance, *patch; /* [io] 1 */
  double* roll_impulse1, *roll_impulse2, *spin_impulse, *tangent_impulse1, *tangent_impulse2; /* [io] 1 */
  double* drift1, *drift2;   /* [io] 1 */
  double* separation;        /* [io] 4 */
  double* point, *shift;     /* [io] 12 */
  double* normal_impulse, *max_normal_impulse, *relative_velocity; /* [io] 4 */
  double* gap;               /* [io] 1 */
  double* now;               /* [out] 4 */
  double* ra, *rb, *la, *lb; /* [out] 12 */
  double* adjusted, *normal_mass; /* [out] 4 */
  double* jac;               /* [out] 48 */
  double* block;             /* [out] 16 */
  double* lever;             /* [out] 4 */
  double* anchor_a, *anchor_b; /* [out] 3 */
  double* fjac;              /* [out] 24 */
  double* tangent_mass;      /* [out] 3 */
  double* roll_mass1, *roll_mass2, *spin_mass; /* [out] 1 */
  uint32_t* pid;             /* [io] 4 */
  /* what the pair list keeps, by pair (src/pairs.js): bm_download writes the entries store()
   * writes (src/contact.js:1067-1088), those of the pairs of solved manifolds with sub < 0 (a
   * pair's points, twist, friction, slip, and ids and impulses for its points); every other
   * entry is left as the caller's arrays hold it. bm_upload does not read them. */
  uint8_t* pair_points;      /* [out] 1 */
  uint32_t* pair_pid;        /* [out] 4 */
  double* pair_normal_impulse; /* [out] 4 */
  double* pair_twist, *pair_friction, *pair_slip; /* [out] 3 */
} bm_columns;

typedef struct bm_context bm_context;

/* A context on HIP device `device`, with room grown as uploads need it. */
int bm_create(bm_context** out, int32_t device);
int bm_destroy(bm_context* ctx);
/* The last error's message, "caller: ..." as the reference words its errors
 * (src/core/check.js:15-22); "" after a success. Owned by the context. */
const char* bm_last_error(const bm_context* ctx);

/* Copies a batch's columns into the context's device memory (HBM). Indices are
 * checked (slots, pairs, point counts); if starts, solved and row describe a
 * layout by colour the batch is ready for bm_soft_step, else bm_colour lays it out. */
int bm_upload(bm_context* ctx, const bm_columns* host);

/* src/graph.js:105-214 and src/contact.js:1318-1329 on the device, for each world of
 * the uploaded batch whose manifolds are in the order narrow() found them: colour
 * each (the reference's greedy lowest-free colour of 32, overflow, resting), lay the
 * rows out colour by colour and by b's slot within a colour, move every column narrow()
 * writes to its row, set row / order / starts / solved, and give resting rows rest()'s
 * values. Bit-identical to the reference's colour(), arrange() and rest(). */
int bm_colour(bm_context* ctx);

/*
 * The soft step on the device, on the uploaded batch: for each world what
 * src/world.js step() runs from inertias() to store() (src/world.js:961-1031)
 * for a world with no joints —
 *   contact.js inertias (1302), the start poses (world.js:964-974),
 *   prepare (1337-1445), makeSoft (1590-1597),
 *   `substeps` times: world._accelerate (1495), warmStart (1604-1663),
 *   solve with bias (1782-2090), world._move with tumble/advance
 *   (1533-1563, 2739-2834), solve relaxing;
 *   restitution (2229-2251), store (1067-1182).
 * Colours run one after another, each colour's manifolds in parallel, the
 * overflow in order (src/graph.js:1-60). Results are the reference's bits.
 */
int bm_soft_step(bm_context* ctx);

/* Copies every [out]/[io] column of the uploaded batch back into `host`. */
int bm_download(bm_context* ctx, bm_columns* host);

/* bm_upload and bm_download for worlds whose columns lie apart: worlds[i] describes one or
 * more worlds with its own offset tables, as bm_upload takes them; the batch is every part's
 * worlds in turn, and bm_download_worlds writes each part's columns back into its own arrays.
 * No bm_upload counterpart in the reference: it lets separately stored worlds share one launch
 * (bm_batch_step). */
int bm_upload_worlds(bm_context* ctx, int32_t n, const bm_columns* const* worlds);
int bm_download_worlds(bm_context* ctx, int32_t n, bm_columns* const* worlds);

/* bm_upload + bm_soft_step + bm_download: what a binding of world.step's
 * contact phase calls with the world's own arrays. */
int bm_soft_step_host(bm_context* ctx, bm_columns* host);

/*
 * The narrowphase on the device, one pair a lane: src/collide.js collide(a, b,
 * margin, tolerance, out) (204-242) for spheres and boxes (349-403, 713-1229), the
 * shapes of pair k given as records a[19k..], b[19k..] = kind (0 sphere, 2 box),
 * centre 3, axes 9 (the shape's x, y, z axes in world space), half 3 (box: half
 * extents; sphere: radius), weight 3 (gravity times gravityScale for a dynamic
 * body, else 0; reducePoints reads it). Out, per pair: count (0-4; -1 for a
 * capsule or hull pair, not handled), normal 3 (a to b), points 12, separations 4,
 * ids 4 (the reference's feature ids), loss (Manifold.loss; may be NULL). Bit-identical
 * to the reference's collide(). Hulls (kind 3) are those of the table bm_set_hulls set.
 */
int bm_collide_host(bm_context* ctx, int32_t n, const double* a, const double* b, const double* margin,
                    const double* tolerance, int32_t* count, double* normal, double* points, double* separations,
                    uint32_t* ids, double* loss);

/*
 * The narrow driver's meet on the device, one pair a lane: src/contact.js meet
 * (627-640) for two whole bodies. collide() as above; when the pair is apart by
 * more than the tolerance and moving() (686-693) says it closes by more within the
 * step, the time of impact (src/impact.js:63-89, conservative advancement over
 * src/gjk.js closest), collide() again where the bodies meet with the margin left,
 * and back() (712-732) carrying the points to the step's start.
 * Body records, 32 numbers: the shape record above (19), rotation 4, velocity 3,
 * spin 3, gravity scale (0 unless dynamic), reach, flags (1 dynamic, 2 spin capped at
 * pi/dt: dynamic and not a plain sphere). Pair records, 6 numbers: margin,
 * tolerance, dt, gravity 3. Out as bm_collide_host, plus shift 12 a pair (b's end
 * from a's at the start, src/contact.js SHIFT) and flags (1 shifted: met at the
 * moment of impact; 2 a time of impact was asked, world._impacts).
 */
int bm_meet_host(bm_context* ctx, int32_t n, const double* a, const double* b, const double* pair, int32_t* count,
                 double* normal, double* points, double* separations, uint32_t* ids, double* shift, int32_t* flags,
                 double* loss);
/* bm_meet_host with each body's record given once: table holds `bodies` records of 32 numbers
 * (the layout bm_meet_host's a and b have), and pair k is the bodies table[ia[k]] and
 * table[ib[k]] with its 6 numbers in pair; results as bm_meet_host's. For a pass whose bodies
 * meet several others each, what crosses to the device is a record a body, not two a pair. */
int bm_meet_indexed(bm_context* ctx, int32_t bodies, const double* table, int32_t n, const int32_t* ia,
                    const int32_t* ib, const double* pair, int32_t* count, double* normal, double* points,
                    double* separations, uint32_t* ids, double* shift, int32_t* flags, double* loss);
/* bm_meet_indexed with the results packed, for a caller that reads the manifolds found only:
 * count[k] and flags[k] for every pair; row[k] the packed row of pair k when count[k] > 0 (rows
 * numbered in pair order), else -1; its normal (3 a row), points (12), separations (4), ids
 * (4), shift (12) and loss (1) in that row. The outputs have room for n rows; only the rows
 * found come back from the device (most pairs of a settled pile whose boxes touch find none). */
int bm_meet_packed(bm_context* ctx, int32_t bodies, const double* table, int32_t n, const int32_t* ia,
                   const int32_t* ib, const double* pair, int32_t* count, int32_t* flags, int32_t* row,
                   double* normal, double* points, double* separations, uint32_t* ids, double* shift, double* loss);

/* The hulls shape records name by number (a hull's record: kind 3, its number in place of
 * half[0]): n hulls in their bodies' frames, as src/hull.js's Hull holds them. Hull h's
 * corners are vertices[3 hull_v[h] .. 3 hull_v[h+1]), its faces' planes (unit outward
 * normal, distance) planes[4 hull_f[h] .. ), its face f's corners loops[hull_l[h] +
 * start[hull_f[h] + f] ..] counter-clockwise from outside. Each edge, its faces, the
 * inscribed ball and the reach are derived as src/hull.js:45-80 derives them. Up to 64
 * faces, 96 edges and 16 corners a face. The context keeps the table until the next call. */
int bm_set_hulls(bm_context* ctx, int32_t n, const int32_t* hull_v, const int32_t* hull_f, const int32_t* hull_l,
                 const double* vertices, const int32_t* start, const int32_t* loops, const double* planes);

/* Heightfields (src/mesh.js:288 Heightfield) that level records name by number: field f has
 * rows[f] rows of columns[f] heights from heights + offsets[f], row by row (x along columns,
 * z along rows, centred on its body), samples spacing[f] apart. The context keeps the table
 * until the next call. */
int bm_set_fields(bm_context* ctx, int32_t n, const int32_t* columns, const int32_t* rows, const double* spacing,
                  const int64_t* offsets, const double* heights);

/*
 * touchLevel (src/triangle.js:1078-1131) on the device, one shape a lane: shape record k (as
 * bm_collide_host's: a sphere, capsule, box or hull) against field field[k], whose body is at
 * position 3k turned by axes 9k (columns), looking at the cells under box 6k (lo 3, hi 3, in
 * the field's frame, as src/contact.js:832-881 level() works it out). Each triangle is touched
 * (src/triangle.js:104 touch: faces, edges and corners, seams and folds from
 * src/mesh.js:409 edge), its points grouped by normal and reduced to four a direction.
 * Out: manifolds[k], how many directions (all of them, even past cap), then cap slots a
 * shape of count, normal 3 (from the shape to the field), points 12, separations 4, ids 4;
 * unused slots zero.
 */
int bm_level_host(bm_context* ctx, int32_t n, const int32_t* field, const double* position, const double* axes,
                  const double* shapes, const double* boxes, const double* margin, const double* tolerance, int32_t cap,
                  int32_t* manifolds, int32_t* count, double* normal, double* points, double* separations,
                  uint32_t* ids);
/* bm_level_host with the results packed by pair, for a caller that reads only the manifolds
 * found: pair k's min(manifolds[k], cap) manifolds are rows start[k] .. start[k + 1] - 1 of
 * count, normal (3 a row), points (12), separations (4) and ids (4), in the order bm_level_host
 * gives them; start has n + 1 entries, the outputs room for n * cap rows. Only the rows found
 * come back from the device, and rows past start[n] are left as they were. */
int bm_level_packed(bm_context* ctx, int32_t n, const int32_t* field, const double* position, const double* axes,
                    const double* shapes, const double* boxes, const double* margin, const double* tolerance,
                    int32_t cap, int32_t* manifolds, int32_t* start, int32_t* count, double* normal, double* points,
                    double* separations, uint32_t* ids);

/*
 * bm_level_host without allocating: for callers that cannot allocate device memory during a
 * call (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference. The same touchLevel, bit for bit, with every array in
 * device memory the caller owns, in bm_level_host's layout; the field and hull tables are
 * the context's (bm_set_fields, bm_set_hulls: set up beforehand). The pass's working memory
 * (each pair's share of the found points) is a fixed 22.5 MB block of device memory the
 * library holds on each device until the process ends: at most 65,536 pairs a call, and at
 * most 262,144 found points a call, 7 for each triangle under each pair's box. A call
 * beyond either, or naming a shape kind, hull or field the tables lack, fails with
 * BM_ERR_ARGUMENT before the pass runs. The call allocates nothing, waits for its work,
 * and then the block holds the last call's contents; two calls that overlap on one device
 * (two contexts on two threads) race on the block.
 */
int bm_level_noalloc(bm_context* ctx, int32_t n, const int32_t* field, const double* position, const double* axes,
                     const double* shapes, const double* boxes, const double* margin, const double* tolerance,
                     int32_t cap, int32_t* manifolds, int32_t* count, double* normal, double* points,
                     double* separations, uint32_t* ids);

/* How bm_soft_step launches, for tuning (bearing_amd/tune.py picks per workload):
 * shape 0 = 256 threads a workgroup with registers for one wave a SIMD, 1 = 256
 * threads and two waves a SIMD, 2 = 512 threads and two waves; resident = 1 keeps
 * each world's body velocities in LDS for the phase when the largest world's fit
 * (48 bytes a body), 0 leaves them in HBM. Defaults: 0, 1. Results are the same
 * bits for every setting. */
int bm_set_launch(bm_context* ctx, int32_t shape, int32_t resident);

/* keep = 1: the columns only the soft step writes and then reads (now, ra, rb, normal_mass,
 * jac, block, lever, anchor_a, anchor_b, fjac, tangent_mass, roll_mass1, roll_mass2,
 * spin_mass) stay on the device: bm_upload does not send them (they may be NULL) and
 * bm_download does not write them back; and the columns the step writes in every row before
 * reading them there (la, lb, adjusted, max_normal_impulse, relative_velocity, gap, sliding)
 * are not sent, only returned. For callers that never read them between steps (the
 * whole-step path); results are the same bits. Default 0. */
int bm_set_device_only(bm_context* ctx, int32_t keep);

/* The phase's device part in one call: bm_colour then bm_soft_step. With the graph on
 * (bm_set_graph) and a batch of more than one world, their launches (the world table, the
 * colouring, arrange and its copy back, rest, the soft step and its two timing events) are
 * captured from the context's stream into a HIP graph and launched as one; the executable graph
 * of the call before is updated to this call's arguments (the arena is laid out by each step's
 * counts, so addresses move) and made anew only when the update is refused. Results are the
 * same bits either way; bm_last_kernel_ms is then the whole graph's time (colouring and soft
 * step), not the soft step's alone. */
int bm_contact_phase(bm_context* ctx);
int bm_set_graph(bm_context* ctx, int32_t on);
/* bm_contact_phase's graphs so far: out[0] launches, out[1] executable graphs made, out[2]
 * updates of the one before. */
int bm_graph_counts(const bm_context* ctx, int64_t* out);

/* version != 0: the bodies' columns no step changes (dims, inertia, inv_inertia, gravity_scale,
 * damping, angular_damping, type, shape, lock, framed, joint_head) are kept on the device in a
 * block of their own, sent by the first upload under this version and not again while uploads
 * keep the same version and number of bodies. The caller gives a new version whenever any of
 * those columns, or the worlds or bodies it uploads, change. 0 (the default) sends them with
 * every upload. Results are the same bits. */
int bm_set_body_version(bm_context* ctx, uint64_t version);

/* When bm_soft_step spreads one world over the whole device: a batch of one world with at
 * least min_manifolds manifolds (a quarter of that if the world's velocities do not fit in one
 * workgroup's LDS) is stepped by every workgroup the device can keep resident, each colour's
 * manifolds spread over all of them with a grid barrier between colours (0: never; default
 * 16384). Results are the same bits either way. */
int bm_set_wide(bm_context* ctx, int32_t min_manifolds);

/* How bm_colour runs: allow = 1 (the default) colours each world in one workgroup's LDS,
 * one manifold after another as the reference does, when every world of the batch fits
 * there (8 bytes a sort key, 13 a manifold, 5 a body); a batch of one world too large for
 * that whose bodies fit (5 bytes a body) is walked the same way with its manifolds streamed
 * from HBM, and its rows laid out by a stable radix sort. allow = 0, or worlds too large for
 * either, take the HBM path, a priority fixed point over each manifold's earlier neighbours.
 * allow = 2 streams a batch of one world whatever its size. All give the reference's colours. */
int bm_set_colour_lds(bm_context* ctx, int32_t allow);

/* Bytes of LDS a soft-step workgroup can hold velocities in on this context's
 * device (a world fits when 48 x its bodies is no more); -1 without a context. */
int64_t bm_lds_limit(bm_context* ctx);

/* Milliseconds the last bm_soft_step's kernel took on the device (HIP events). */
double bm_last_kernel_ms(const bm_context* ctx);

/* ---- the broadphase's data-parallel passes (src/world.js:1574-1630, src/broadphase.js) ----
 *
 * The reference's Broadphase (src/broadphase.js:56-219) re-files a body in its tree when its
 * bounds leave its fat box, walks both trees for every body it marked, and hands what the walks
 * meet to the pair list in that order; the narrow pass first drops pairs whose fat boxes no
 * longer overlap (src/pairs.js:340-356 holds). These calls do the per-body and per-pair parts
 * on the device, for one world at a time; the host keeps the trees' shape (src/tree.js insert,
 * move, rotate) and takes the results in the reference's order. A context holds one world's
 * bodies and trees for them (the last ones given). */

/* Every body's constants and state for the bounds, by slot (src/world.js:1580-1630): type,
 * shape, dims 3, gravityScale, reach, size, local 6 (a hull's or heightfield's box in its own
 * frame: centre 3, half extents 3; anything for other shapes), position 3, rotation 4,
 * velocity 3, angular 3. */
typedef struct bm_broad_bodies {
  int32_t n;
  const uint8_t *type, *shape;
  const double *dims, *gravity_scale, *reach, *size, *local;
  const double *position, *rotation, *velocity, *angular;
} bm_broad_bodies;
int bm_broad_set_bodies(bm_context* ctx, const bm_broad_bodies* bodies);
/* The state of n bodies (slots[i]) changed on the host: their position, rotation, velocity and
 * angular, read from the per-slot host arrays, replace the device's. */
int bm_broad_set_state(bm_context* ctx, int32_t n, const int32_t* slots, const double* position,
                       const double* rotation, const double* velocity, const double* angular);
/* After a soft step of a batch of one world with as many bodies as bm_broad_set_bodies gave:
 * the step's positions, rotations, velocities and spins become the broadphase's, on the device
 * (no copy through the host). */
int bm_broad_keep(bm_context* ctx);
/* The two trees (static, moving; src/broadphase.js:56-70) as src/tree.js keeps them: nodes
 * (capacity), root (-1 empty), box 6 a node, child1, child2, height (0 a leaf), item (a leaf's
 * body slot); and each body's leaf (leaf, by slot) and tree (which: 0 static, 1 moving, -1 none).
 * The trees stay until the next call. */
typedef struct bm_broad_tree {
  int32_t nodes, root;
  const double* box;
  const int32_t *child1, *child2, *height, *item;
} bm_broad_tree;
int bm_broad_set_trees(bm_context* ctx, const bm_broad_tree* trees, int32_t n, const int32_t* leaf,
                       const int8_t* which);
/* src/world.js:1574-1578 (_bounds) and the test of src/broadphase.js:140-159 (update) for the n
 * awake bodies list[0..n), stepping dt under gravity of length gravity_length: out (16 n,
 * field-major: out[k * n + i]) each body's axes 9, motion and box 6; refile[i] = 1 where the
 * body is in the moving tree and the box left its leaf's fat box. */
int bm_broad_bounds(bm_context* ctx, int32_t n, const int32_t* list, double dt, double gravity_length, double* out,
                    uint8_t* refile);
/* src/broadphase.js:161-179 (candidates): for each of the n marked bodies moved[0..n), its leaf's
 * box walked down the static tree (unless the body is static, by the type bm_broad_set_bodies
 * gave), then the moving tree (src/tree.js
 * query, child1 first): counts[i] bodies met other than itself, written from hits[starts[i]] in
 * walk order. *total: every body's count together; hits is written only when that is at most
 * capacity (call again with room for *total). */
int bm_broad_query(bm_context* ctx, int32_t n, const int32_t* moved, int32_t* counts, int32_t* starts, int32_t* hits,
                   int64_t capacity, int64_t* total);
/* src/pairs.js:340-356 (holds): apart[i] = 1 where the fat boxes of pair (a[i], b[i])'s leaves
 * no longer overlap (src/broadphase.js:181-187). */
int bm_broad_holds(bm_context* ctx, int32_t n, const int32_t* a, const int32_t* b, uint8_t* apart);

/* ---- a whole world, stepped as src/world.js step (880-1072) steps it ----
 *
 * For worlds of dynamic and static boxes, spheres, capsules and hulls, and static heightfields,
 * with no joints, sensors,
 * characters or compounds. The host keeps the reference's order-defining serial
 * state (the broadphase trees, the pair list, the narrow driver's order, the
 * sleeping islands); the device runs each step's narrowphase batches, colouring
 * and soft step. Transforms match the reference's bit for bit. A step that would
 * need what this path does not handle fails with BM_ERR_UNSUPPORTED and a message,
 * rather than going on differently. */
typedef struct bm_world bm_world;

/* Bodies as addBody resolves them (src/world.js:563-578, _write 2152-2273), one
 * entry a body: the columns the reference keeps by slot. type: 0 dynamic,
 * 1 static; shape: 0 box (dims = full size), 1 sphere (dims[0] = radius),
 * 2 capsule (dims[0] = radius, dims[1] = height end to end), 3 hull (hull = its
 * number in the table bm_world_set_hulls set, in the frame below), 6 heightfield (static;
 * field = its number in the table bm_world_set_fields set).
 * position and rotation are the step's (world._position, _rotation): a framed body's
 * centre of mass and principal axes. framed, offset (3) and frame (4) are
 * world._framed, _offset and _frame (src/world.js:2199-2205): the transforms give a
 * framed body's own origin and turn back as src/world.js:1917-1930 (_pose) does. A
 * framed box, sphere or capsule is a body with parts, which this world does not take.
 * hull may be NULL when no body is a hull, field when none is a heightfield; framed, offset
 * and frame when none is framed. */
typedef struct bm_bodies {
  int32_t n;
  const uint8_t *type, *shape, *lock;
  const double *position, *rotation, *velocity, *angular, *dims, *inertia, *inv_inertia; /* 3, 4, 3, 3, 3, 3, 3 */
  const double *mass, *inv_mass, *gravity_scale, *damping, *angular_damping, *size, *reach, *friction,
      *restitution;
  const int32_t* hull;
  const uint8_t* framed;
  const double *offset, *frame; /* 3, 4 */
  const int32_t* field;
} bm_bodies;

/* What bm_world_read copies out; any pointer may be NULL. transforms: 7 floats a
 * slot (position, rotation), as the reference's Float32Array world.transforms.
 * rows, soft_bytes and soft_ms describe the last step's solve, for measurement (no
 * counterpart in the reference): the points of its solved manifolds (a step is 2 x
 * substeps iterations over each), the bytes its sweeps must read and write at the least
 * (each field once a pass, each body's fields once a manifold end; the sum over manifolds
 * of n points of S warm starts of 201 + 16 [n <= 2] + 56 n + 304 bytes, S - 1 solves of
 * F + n (152 + 8 [n = 1]) + B and S + 1 relaxes of F - 24 + n (208 + 8 [n = 1]) + B + 160,
 * F = 394 + 56 [n <= 2] + 8 n^2 [n >= 2], B = 208 + 96 [both ends move] + 48 [one is
 * static]), and the soft-step kernel's device time (a batch's worlds share one launch and
 * report its time). */
typedef struct bm_world_state {
  int32_t top, awake, contacts, pairs, frozen;
  int64_t impacts, catches; /* world._impacts, world._catches */
  float* transforms;
  double *position, *rotation, *velocity, *angular;
  int64_t rows;
  double soft_bytes, soft_ms;
  float* previous; /* world.previous (src/world.js:547-556): each slot's transform as its last
                      step began (a body that fell asleep: as the step after began) */
} bm_world_state;

/* new World({gravity, substeps, sleeping}) (src/world.js:314): gravity may be NULL
 * for (0, -9.81, 0); substeps 0 for 8. */
int bm_world_create(bm_world** out, int32_t device, const double* gravity, int32_t substeps, int32_t sleeping);
int bm_world_destroy(bm_world* world);
/* The last error's message, "" after a call that succeeded. */
const char* bm_world_error(const bm_world* world);
/* addBody (src/world.js:563) for each body, slots in order. */
int bm_world_add(bm_world* world, const bm_bodies* bodies);
/* The hulls the world's hull bodies name (their world._inner hulls, src/hull.js Hull),
 * in the layout and with the limits of bm_set_hulls. Replaces the table; fails if a body
 * already added names a hull the new table lacks. */
int bm_world_set_hulls(bm_world* world, int32_t n, const int32_t* hull_v, const int32_t* hull_f,
                       const int32_t* hull_l, const double* vertices, const int32_t* start, const int32_t* loops,
                       const double* planes);
/* The heightfields the world's heightfield bodies name, in the layout of bm_set_fields
 * (src/mesh.js:288 Heightfield). Replaces the table; fails if a body already added names a
 * field the new table lacks. */
int bm_world_set_fields(bm_world* world, int32_t n, const int32_t* columns, const int32_t* rows,
                        const double* spacing, const int64_t* offsets, const double* heights);
/* world.step(dt) (src/world.js:880). A dt that is not a positive finite number is refused
 * with the world unchanged (src/world.js:881-884). A step that fails after it began (a sweep
 * this path does not handle, a device error) leaves the world part-way through that step:
 * every later step of the world returns the same code and message. */
int bm_world_step(bm_world* world, double dt);
int bm_world_read(const bm_world* world, bm_world_state* out);
/* world.moved(since) (src/world.js:771-803): into out (room for every slot) the slots of the
 * bodies that moved in any step after step `since`, and always of those that moved in the last
 * one, with bodies added since the last step; *count of them. The order is the reference's:
 * the awake list's, then the moved log's from its newest entry, or, for a `since` older than
 * the log keeps, every slot in order (src/world.js:1465-1492). A body is stamped when it moves
 * awake and unfrozen in a step (src/world.js:1040-1057), when it is added (1450-1454), and in
 * the step after it fell asleep (912-918). */
int bm_world_moved(bm_world* world, int32_t since, int32_t* out, int32_t* count);
/* The last step's contact events (src/world.js:690-723 contacts(), src/events.js): which pairs
 * began touching in it and which stopped, in the order the step wrote them (ends as pairs were
 * let go in the narrow pass, src/pairs.js:279-284; then, in the order the narrow pass found the
 * pairs, a begin or an end for each pair whose touching changed, src/contact.js:2178-2213).
 * Bodies by slot (the reference's ids are generation * 2^24 + slot). A begin carries the
 * deepest point moved on by half of what the two bodies moved in the step, the normal from a
 * to b and the fastest closing speed among the points; an end's point, normal and speed are 0.
 * The pointers are the world's own and hold until its next step or bm_world_destroy. */
/* From how many pairs a world stepped alone (bm_world_step) takes a narrow pass's work, and
 * from how many awake bodies or contact rows its other large loops, to its host threads
 * (default 8,192). No counterpart in the reference, whose workers take a pass's pairs from a
 * fixed size (src/world.js:186); the results are the same bits either way, so this only
 * moves work between the calling thread and the world's threads. */
int bm_world_set_parallel(bm_world* w, int32_t from);
/* Where a world's broadphase passes run (src/world.js:1574-1630 bounds, src/broadphase.js:140-187
 * update and candidates, src/pairs.js:340-356 holds): BM_BROAD_HOST on the host's threads;
 * BM_BROAD_DEVICE on the device (bm_broad_*), the host re-filing bodies in the trees and taking
 * what the walks met in the reference's order; BM_BROAD_AUTO (the default) on the device for a
 * world with its own context and at least 4,096 awake bodies. A batch's worlds take the host's.
 * Results are the same bits every way. No counterpart in the reference. */
#define BM_BROAD_HOST 0
#define BM_BROAD_DEVICE 1
#define BM_BROAD_AUTO 2
int bm_world_set_broadphase(bm_world* world, int32_t mode);
#define BM_EVENT_BEGIN 0
#define BM_EVENT_END 1
typedef struct bm_world_events {
  int32_t count;
  const uint8_t* type;       /* BM_EVENT_BEGIN or BM_EVENT_END */
  const int32_t *a, *b;      /* slots, a < b */
  const double* point;       /* 3 an event */
  const double* normal;      /* 3 an event */
  const double* speed;
} bm_world_events;
int bm_world_events_of(const bm_world* world, bm_world_events* out);
/* Milliseconds the world's steps have spent so far in each part, into out[BM_WORLD_TIMES]:
 * 0 bounds, broadphase and preparing the first narrow pass; 1 the narrowphase on the device
 * (with its copies); 2 taking the narrowphase's results in order and the later passes'
 * preparing; 3 freezing; 4 the contact phase on the device (upload, colour, soft step,
 * download); 5 the catch, touched, transforms and sleep. */
#define BM_WORLD_TIMES 6
int bm_world_times(const bm_world* world, double* out);
/* Shows `tap` the columns each later step hands the device for its contact phase, before the
 * device sees them: bm_columns of this world alone (one world's offset tables), as bm_upload
 * takes them, the manifolds in the order the narrow pass left them (before colouring). The
 * pointers hold only for the call, which runs on the stepping thread. For replaying one step's
 * soft phase elsewhere: another context, a CPU restatement, a profiler. NULL stops it. No
 * counterpart in the reference: for measurement and tests. */
typedef void (*bm_columns_tap)(void* user, const bm_columns* columns);
int bm_world_tap(bm_world* world, bm_columns_tap tap, void* user);

/* ---- worlds stepped together (the batched form) ----
 *
 * A batch steps its worlds as bm_world_step steps each one, with the same results bit for
 * bit, sharing the device between them: each world's order-defining host work (trees, pair
 * list, narrow order, freezing, catch, sleep) runs on one of `threads` host threads (0 or 1:
 * the calling thread alone), every narrowphase pass of every world goes to the device in one
 * launch, and every world's colouring and soft step in one launch each. No counterpart in the
 * reference, which steps one world at a time. */
typedef struct bm_batch bm_batch;
int bm_batch_create(bm_batch** out, int32_t device, int32_t threads);
/* Destroys the batch and every world it made. */
int bm_batch_destroy(bm_batch* batch);
const char* bm_batch_error(const bm_batch* batch);
/* A new world of the batch, as bm_world_create makes one (bodies, hulls and fields are added
 * to it as to any world); it lives until bm_batch_destroy. */
int bm_batch_world(bm_batch* batch, bm_world** out, const double* gravity, int32_t substeps, int32_t sleeping);
/* Whether the batch's contact phase goes to the device as a HIP graph (bm_contact_phase);
 * default on. */
int bm_batch_set_graph(bm_batch* batch, int32_t on);
/* The batch's soft-step launch shape and velocity residency, as bm_set_launch sets a context's
 * (an autotuner times them on the batch's own steps); every world's bits are the same for each. */
int bm_batch_set_launch(bm_batch* batch, int32_t shape, int32_t resident);
int bm_batch_graph_counts(const bm_batch* batch, int64_t* out);
/* Every world of the batch, one step of dt. A step that fails after the worlds began theirs
 * leaves all of them part-way: every later step of the batch returns the same code and
 * message, and so does the message of each of its worlds. */
int bm_batch_step(bm_batch* batch, double dt);
/* As bm_world_times, for the batch's steps (host parts on the batch's threads, wall time). */
int bm_batch_times(const bm_batch* batch, double* out);

/* ---- The adjoint: gradients of a recorded trajectory ----------------------------------
 *
 * No counterpart in the reference, which has no derivatives. A world records the steps it
 * takes (bm_world_record); each step's solved manifolds are kept held on their bodies, as
 * the reference's REUSE keeps a manifold (src/contact.js:388-508): each point's end on a in
 * a's frame and on b in b's frame, the normal in a's frame, the feature ids, the colours and
 * the order the step found them in. A held trajectory (bm_adjoint_create) steps the recorded
 * worlds again from their first recorded state with every manifold placed from where its
 * bodies are (the REUSE rebuild), the impulses carried as the pairs carried them
 * (src/contact.js:999-1036), and the reference's soft step (src/world.js:961-1031); at the
 * recorded parameters it follows the recorded steps (bit for bit where the reference itself
 * reused the manifold; where the reference found one afresh, up to the rounding of the
 * rebuild; a manifold found at the moment of impact keeps its carry back as a constant).
 * bm_adjoint_backward gives the gradient of any function of the final state with
 * respect to the initial state and the bodies' mass, friction and restitution, on the
 * forward pass's active set (Murty sets, friction cone, clamps, bounces): which manifolds a
 * step has is held, where their points lie moves with the bodies.
 *
 * The forward pass keeps every body's state every `every` steps and every step's stored
 * impulses; the backward pass steps each segment again from its first state writing a tape
 * and walks it backwards (reverse colour sweeps, the integrator's adjoint), one workgroup a
 * world. The recorded world must step with sleeping off (so nothing freezes or sleeps),
 * without catches, heightfields or bodies added while recording; it may step alone or in a
 * batch (bm_batch_step records each recording world of the batch).
 *
 * State per body slot: p 3, q 4, v 3, w 3 (BM_HELD_STATE doubles); slots of the worlds one
 * after another, each world's as its own slots number them. */
#define BM_HELD_STATE 13
/* Starts (on != 0) or stops recording. Starting clears what was recorded. */
int bm_world_record(bm_world* world, int32_t on);
/* Steps recorded so far. */
int bm_world_recorded(const bm_world* world);
typedef struct bm_adjoint bm_adjoint;
/* The held trajectory of n recorded worlds (each from its first recorded step to its last)
 * on a device, keeping state every `every` steps. The worlds may be destroyed afterwards. */
int bm_adjoint_create(bm_adjoint** out, int32_t device, bm_world* const* worlds, int32_t n, int32_t every);
int bm_adjoint_destroy(bm_adjoint* a);
const char* bm_adjoint_error(const bm_adjoint* a);
/* Body slots (the worlds' together) and steps. */
int bm_adjoint_size(const bm_adjoint* a, int32_t* bodies, int32_t* steps);
/* The recorded values: initial state (BM_HELD_STATE a slot), mass, friction, restitution. */
int bm_adjoint_recorded(const bm_adjoint* a, double* state0, double* mass, double* friction, double* restitution);
/* The forward pass from state0 with the given parameters; the final state into state_out. */
int bm_adjoint_forward(bm_adjoint* a, const double* state0, const double* mass, const double* friction,
                       const double* restitution, double* state_out);
/* After bm_adjoint_forward: from the final state's adjoint (dL/dstate), the adjoints of the
 * initial state and of every slot's mass, friction and restitution (0 for static bodies'
 * mass). */
int bm_adjoint_backward(bm_adjoint* a, const double* state_bar, double* state0_bar, double* mass_bar,
                        double* friction_bar, double* restitution_bar);
/* Milliseconds the last forward and backward pass took on the device. */
int bm_adjoint_times(const bm_adjoint* a, double* forward_ms, double* backward_ms);
/* After bm_adjoint_forward: for each world, a 64-bit record of the branches its forward pass
 * took — its active set: each point's anchor weight case, each solve's target cases, Murty
 * set, friction bias and cone, roll and spin clamps, each bounce taken or not and clamped,
 * each spin cap and the tumble's branches, in the order taken. Two forward passes of a world
 * took the same branches when its records are equal; the gradient is that of the pass's
 * branches, so finite differences agree with it only between inputs whose records agree. */
int bm_adjoint_branches(bm_adjoint* a, uint64_t* per_world);

/* The held trajectory as the device takes it (bm_held_load); host pointers. Rows are every
 * step's solved manifolds, step after step and within a step world after world, after
 * `first` rows that only hold what the worlds' pairs kept before recording began. */
typedef struct bm_held {
  int32_t worlds, steps, every, bodies, rows, first;
  const int32_t* body_offset; /* worlds + 1 */
  const int32_t* row_offset;  /* steps * worlds + 1 */
  const int32_t* starts;      /* BM_STARTS per (step, world), relative to its first row */
  const int32_t* narrow;      /* per row: its (step, world)'s rows in narrow order, relative */
  const int32_t* substeps;    /* per world */
  const double* dt;           /* per world */
  const double* gravity;      /* 3 per world */
  /* per body */
  const uint8_t* type, *shape, *tumbles, *capped;
  const double* rounded, *segment, *gravity_scale, *damping, *angular_damping, *mass0, *inv_mass0, *inertia0,
      *inv_inertia0;
  const double* left, *ceiling; /* the tumble's own state at the start: 3, 1 */
  /* per row */
  const int32_t* a, *b, *src;
  const uint8_t* count;
  const uint32_t* pid;        /* 4 */
  const double* local_normal; /* 3 */
  const double* ends;         /* 24: per point a's end 3, b's end 3 */
  const double* shift;        /* 12: per point, b's carry back from the moment of impact */
  const double* tolerance, *patch;
  /* the stores of the first rows: impulses 4, friction 3, twist 3, slip 3 */
  const double* store_impulse, *store_friction, *store_twist, *store_slip;
} bm_held;
int bm_held_load(bm_context* ctx, const bm_held* held);
/* state0: BM_HELD_STATE a body; the rest one a body. */
int bm_held_forward(bm_context* ctx, const double* state0, const double* mass, const double* friction,
                    const double* restitution, double* state_out);
int bm_held_backward(bm_context* ctx, const double* state_bar, double* state0_bar, double* mass_bar,
                     double* friction_bar, double* restitution_bar);
int bm_held_times(const bm_context* ctx, double* forward_ms, double* backward_ms);
/* The held trajectory as bm_adjoint_create gave it to the device (bm_held_load's layout): its
 * pointers are into the adjoint's own copy, good until bm_adjoint_destroy. No counterpart in the
 * reference; for checking the held step against another implementation of it. */
int bm_adjoint_held(const bm_adjoint* a, bm_held* out);
/* bm_adjoint_branches for the held trajectory bm_held_load loaded. */
int bm_held_branches(bm_context* ctx, uint64_t* per_world);

#ifdef __cplusplus
}
#endif

#endif /* BEARING_MI355X_H */
