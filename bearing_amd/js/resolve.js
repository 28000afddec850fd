'use strict';
/*
 * addBody's options resolved into the columns the world keeps by slot (bm_bodies):
 * src/world.js:206-216 (DEFAULTS: a dynamic 1 m box, density 1000, friction 0.5,
 * restitution 0, no damping, gravity scale 1), src/world.js:2331-2400 (resolve: the checks
 * this front-end makes) and src/world.js:2152-2273 (_write: mass from density and volume
 * unless a mass is given, the principal inertias, the thinnest size and the reach, inverses
 * where positive; a static body has no inverse mass). Boxes, spheres and capsules; static
 * heightfields (src/mesh.js:288-305); and hulls given as src/hull.js build() makes them
 * (corners, faces, loops and planes in the frame their points were given in), whose mass,
 * centre, principal axes and frame this file works out as src/shapes.js and
 * src/world.js:2544-2549 (solidMass) do, and which it moves to their centre of mass and
 * principal axes as src/hull.js:113-134 (moved) does. Plain JavaScript, no addon: it runs
 * without a GPU.
 */

const DEFAULTS = Object.freeze({
  type: 'dynamic', shape: 'box', size: [1, 1, 1], radius: 0.5, height: 2,
  position: [0, 0, 0], rotation: [0, 0, 0, 1], velocity: [0, 0, 0], angularVelocity: [0, 0, 0],
  density: 1000, friction: 0.5, restitution: 0, damping: 0, angularDamping: 0, gravityScale: 1,
});

/* Options this front-end takes; any other is refused rather than ignored. */
const TAKEN = new Set(Object.keys(DEFAULTS).concat(['mass', 'hull', 'heights', 'columns', 'spacing']));
/* The shapes' numbers in bm_bodies (src/world.js:62 SHAPES). */
const SHAPE = { box: 0, sphere: 1, capsule: 2, hull: 3, heightfield: 6 };
/* A double's rounding, 2^-52 (src/shapes.js:188). */
const EPSILON = 1 / 4503599627370496;

function fail(caller, message) {
  throw new Error(`${caller}: ${message}`);
}

function vector(caller, name, v, n) {
  if (!Array.isArray(v) && !ArrayBuffer.isView(v)) fail(caller, `${name} must be an array of ${n} numbers`);
  if (v.length !== n) fail(caller, `${name} must have ${n} numbers, got ${v.length}`);
  for (let i = 0; i < n; i++) {
    if (typeof v[i] !== 'number' || !Number.isFinite(v[i])) fail(caller, `${name} must hold finite numbers`);
  }
  return v;
}

/*
 * A rotation as the world keeps it (src/core/check.js:105-116 quat): scaled by its largest
 * magnitude, then to unit length; all zero is refused.
 */
function rotation(caller, q) {
  vector(caller, 'rotation', q, 4);
  const m = Math.max(Math.abs(q[0]), Math.abs(q[1]), Math.abs(q[2]), Math.abs(q[3]));
  if (!(m > 0)) fail(caller, 'rotation must be 4 finite numbers, not all zero');
  const x = q[0] / m, y = q[1] / m, z = q[2] / m, w = q[3] / m;
  const k = 1 / Math.sqrt(x * x + y * y + z * z + w * w);
  return [x * k, y * k, z * k, w * k];
}

function positive(caller, name, x) {
  if (typeof x !== 'number' || !(x > 0) || !Number.isFinite(x)) fail(caller, `${name} must be a number above 0, got ${x}`);
  return x;
}

function whole(caller, name, x, least) {
  if (!Number.isInteger(x) || x < least) fail(caller, `${name} must be a whole number, at least ${least}, got ${x}`);
  return x;
}

function numbers(caller, name, v, Kind) {
  if (!Array.isArray(v) && !ArrayBuffer.isView(v)) fail(caller, `${name} must be an array of numbers`);
  const out = Kind.from(v);
  for (let i = 0; i < out.length; i++) {
    if (!Number.isFinite(out[i]) || out[i] !== Number(v[i])) fail(caller, `${name} must hold finite numbers`);
  }
  return out;
}

/* src/core/math.js:230-238 (quatMultiply): a (x) b, b's turn then a's. */
function quatMultiply(a, b) {
  const ax = a[0], ay = a[1], az = a[2], aw = a[3], bx = b[0], by = b[1], bz = b[2], bw = b[3];
  return [aw * bx + ax * bw + ay * bz - az * by, aw * by + ay * bw + az * bx - ax * bz,
    aw * bz + az * bw + ax * by - ay * bx, aw * bw - ax * bx - ay * by - az * bz];
}

/* src/core/math.js:211-219 (rotate), not inverted. */
function rotate(q, x, y, z) {
  const ux = q[0], uy = q[1], uz = q[2], w = q[3];
  const tx = 2 * (uy * z - uz * y), ty = 2 * (uz * x - ux * z), tz = 2 * (ux * y - uy * x);
  return [x + w * tx + (uy * tz - uz * ty), y + w * ty + (uz * tx - ux * tz), z + w * tz + (ux * ty - uy * tx)];
}

/* src/shapes.js:136-143 (matrixOf): unit quaternion q's rotation, 9 numbers row by row. */
function matrixOf(q) {
  const x = q[0], y = q[1], z = q[2], w = q[3];
  return [
    1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
    2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
    2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
  ];
}

/* src/world.js:2604-2606 (transposed). */
function transposed(R) {
  return [R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]];
}

/*
 * src/shapes.js:71-106 (hullMass): volume, centre of mass and inertia tensor (xx xy xz yy yz
 * zz) about it, for a density of 1; each face fanned into triangles from its first corner,
 * each closed into a tetrahedron with the corners' mean.
 */
function hullMass(h) {
  const v = h.vertices, V = v.length / 3;
  let rx = 0, ry = 0, rz = 0;
  for (let i = 0; i < V; i++) { rx += v[3 * i]; ry += v[3 * i + 1]; rz += v[3 * i + 2]; }
  rx /= V; ry /= V; rz /= V;
  let vol = 0, cx = 0, cy = 0, cz = 0;
  const C = [0, 0, 0, 0, 0, 0];
  const start = h.start, loops = h.loops;
  for (let f = 0; f + 1 < start.length; f++) {
    const o = loops[start[f]];
    for (let k = start[f] + 1; k + 1 < start[f + 1]; k++) {
      const i = loops[k], j = loops[k + 1];
      const ax = v[3 * o] - rx, ay = v[3 * o + 1] - ry, az = v[3 * o + 2] - rz;
      const bx = v[3 * i] - rx, by = v[3 * i + 1] - ry, bz = v[3 * i + 2] - rz;
      const qx = v[3 * j] - rx, qy = v[3 * j + 1] - ry, qz = v[3 * j + 2] - rz;
      const det = ax * (by * qz - bz * qy) + ay * (bz * qx - bx * qz) + az * (bx * qy - by * qx);
      vol += det / 6;
      const sx = ax + bx + qx, sy = ay + by + qy, sz = az + bz + qz;
      cx += det / 24 * sx; cy += det / 24 * sy; cz += det / 24 * sz;
      const k6 = det / 120;
      C[0] += k6 * (ax * ax + bx * bx + qx * qx + sx * sx);
      C[1] += k6 * (ax * ay + bx * by + qx * qy + sx * sy);
      C[2] += k6 * (ax * az + bx * bz + qx * qz + sx * sz);
      C[3] += k6 * (ay * ay + by * by + qy * qy + sy * sy);
      C[4] += k6 * (ay * az + by * bz + qy * qz + sy * sz);
      C[5] += k6 * (az * az + bz * bz + qz * qz + sz * sz);
    }
  }
  cx /= vol; cy /= vol; cz /= vol;
  C[0] -= vol * cx * cx; C[1] -= vol * cx * cy; C[2] -= vol * cx * cz;
  C[3] -= vol * cy * cy; C[4] -= vol * cy * cz; C[5] -= vol * cz * cz;
  // src/shapes.js:109-112 (inertiaOf): tr(C) I - C.
  const t = C[0] + C[3] + C[5];
  return { volume: vol, center: [rx + cx, ry + cy, rz + cz], tensor: [t - C[0], -C[1], -C[2], t - C[3], -C[4], t - C[5]] };
}

/* src/shapes.js:195-213 (quaternionOf): rotation matrix R (rows) as a unit quaternion, Shepperd's method. */
function quaternionOf(R) {
  const t = R[0][0] + R[1][1] + R[2][2];
  let x, y, z, w;
  if (t >= R[0][0] && t >= R[1][1] && t >= R[2][2]) {
    const s = 2 * Math.sqrt(1 + t);
    w = s / 4; x = (R[2][1] - R[1][2]) / s; y = (R[0][2] - R[2][0]) / s; z = (R[1][0] - R[0][1]) / s;
  } else if (R[0][0] >= R[1][1] && R[0][0] >= R[2][2]) {
    const s = 2 * Math.sqrt(1 + R[0][0] - R[1][1] - R[2][2]);
    x = s / 4; w = (R[2][1] - R[1][2]) / s; y = (R[0][1] + R[1][0]) / s; z = (R[0][2] + R[2][0]) / s;
  } else if (R[1][1] >= R[2][2]) {
    const s = 2 * Math.sqrt(1 + R[1][1] - R[0][0] - R[2][2]);
    y = s / 4; w = (R[0][2] - R[2][0]) / s; x = (R[0][1] + R[1][0]) / s; z = (R[1][2] + R[2][1]) / s;
  } else {
    const s = 2 * Math.sqrt(1 + R[2][2] - R[0][0] - R[1][1]);
    z = s / 4; w = (R[1][0] - R[0][1]) / s; x = (R[0][2] + R[2][0]) / s; y = (R[1][2] + R[2][1]) / s;
  }
  const l = Math.sqrt(x * x + y * y + z * z + w * w);
  return [x / l, y / l, z / l, w / l];
}

/*
 * src/shapes.js:157-185 (principal): the principal moments of tensor T and the unit
 * quaternion whose rotation's columns are their axes, by Jacobi's method: sweeps over the
 * three off-diagonal entries, each cleared by the turn whose tangent Numerical Recipes 11.1
 * gives, until every entry is under a double's rounding of the two moments it joins.
 */
function principal(T) {
  const A = [[T[0], T[1], T[2]], [T[1], T[3], T[4]], [T[2], T[4], T[5]]];
  const R = [[1, 0, 0], [0, 1, 0], [0, 0, 1]];
  const PAIRS = [[0, 1], [0, 2], [1, 2]];
  for (let sweep = 0; sweep < 64; sweep++) {
    let turned = false;
    for (const [p, q] of PAIRS) {
      const apq = A[p][q];
      if (!(Math.abs(apq) > EPSILON * (Math.abs(A[p][p]) + Math.abs(A[q][q])))) { A[p][q] = A[q][p] = 0; continue; }
      turned = true;
      const theta = (A[q][q] - A[p][p]) / (2 * apq);
      const t = (theta >= 0 ? 1 : -1) / (Math.abs(theta) + Math.sqrt(theta * theta + 1));
      const c = 1 / Math.sqrt(t * t + 1), sn = t * c;
      for (let k = 0; k < 3; k++) {
        const akp = A[k][p], akq = A[k][q];
        A[k][p] = c * akp - sn * akq; A[k][q] = sn * akp + c * akq;
      }
      for (let k = 0; k < 3; k++) {
        const apk = A[p][k], aqk = A[q][k];
        A[p][k] = c * apk - sn * aqk; A[q][k] = sn * apk + c * aqk;
      }
      for (let k = 0; k < 3; k++) {
        const rkp = R[k][p], rkq = R[k][q];
        R[k][p] = c * rkp - sn * rkq; R[k][q] = sn * rkp + c * rkq;
      }
    }
    if (!turned) break;
  }
  return { moments: [A[0][0], A[1][1], A[2][2]], rotation: quaternionOf(R) };
}

/*
 * src/hull.js:113-134 (moved): hull h in the frame with origin c turned by r (9 numbers, row
 * by row), each corner r (x - c), each plane's normal r n and distance d - n.c; with the
 * moved hull's reach (its farthest corner from the origin) and size (its least width across
 * a face), as src/hull.js:62-84 derives them.
 */
function moved(h, c, r) {
  const v = h.vertices, V = v.length / 3, out = new Float64Array(3 * V);
  for (let i = 0; i < V; i++) {
    const x = v[3 * i] - c[0], y = v[3 * i + 1] - c[1], z = v[3 * i + 2] - c[2];
    out[3 * i] = r[0] * x + r[1] * y + r[2] * z;
    out[3 * i + 1] = r[3] * x + r[4] * y + r[5] * z;
    out[3 * i + 2] = r[6] * x + r[7] * y + r[8] * z;
  }
  const p = h.planes, F = p.length / 4, planes = new Float64Array(4 * F);
  for (let f = 0; f < F; f++) {
    const nx = p[4 * f], ny = p[4 * f + 1], nz = p[4 * f + 2];
    planes[4 * f] = r[0] * nx + r[1] * ny + r[2] * nz;
    planes[4 * f + 1] = r[3] * nx + r[4] * ny + r[5] * nz;
    planes[4 * f + 2] = r[6] * nx + r[7] * ny + r[8] * nz;
    planes[4 * f + 3] = p[4 * f + 3] - (nx * c[0] + ny * c[1] + nz * c[2]);
  }
  let reach = 0;
  for (let i = 0; i < V; i++) {
    const x = out[3 * i], y = out[3 * i + 1], z = out[3 * i + 2];
    reach = Math.max(reach, Math.sqrt(x * x + y * y + z * z));
  }
  let size = Infinity;
  for (let f = 0; f < F; f++) {
    const nx = planes[4 * f], ny = planes[4 * f + 1], nz = planes[4 * f + 2];
    let low = Infinity;
    for (let i = 0; i < V; i++) low = Math.min(low, nx * out[3 * i] + ny * out[3 * i + 1] + nz * out[3 * i + 2]);
    size = Math.min(size, planes[4 * f + 3] - low);
  }
  return { vertices: out, start: h.start, loops: h.loops, planes, reach, size };
}

/* A hull as build() makes it: corners, faces' first loop entries (one more than the faces), loops, planes. */
function builtHull(caller, h) {
  if (h === null || typeof h !== 'object') fail(caller, 'hull must be a built hull: {vertices, start, loops, planes}');
  const vertices = numbers(caller, 'hull.vertices', h.vertices, Float64Array);
  const start = numbers(caller, 'hull.start', h.start, Int32Array);
  const loops = numbers(caller, 'hull.loops', h.loops, Int32Array);
  const planes = numbers(caller, 'hull.planes', h.planes, Float64Array);
  const V = vertices.length / 3, F = start.length - 1;
  if (V < 4 || vertices.length % 3 !== 0) fail(caller, 'hull.vertices must hold at least 4 corners, 3 numbers each');
  if (F < 4 || planes.length !== 4 * F) fail(caller, 'hull.planes must hold 4 numbers for each face of hull.start');
  if (start[0] !== 0 || start[F] !== loops.length) fail(caller, 'hull.start must run from 0 to the length of hull.loops');
  for (let f = 0; f < F; f++) if (start[f + 1] - start[f] < 3) fail(caller, 'a hull face must have 3 corners or more');
  for (let i = 0; i < loops.length; i++) if (loops[i] < 0 || loops[i] >= V) fail(caller, 'hull.loops must name corners');
  return { vertices, start, loops, planes };
}

/*
 * The columns of `list` (addBody options, one entry a body), as the addon's add() takes
 * them: {n, type, shape, lock, position, rotation, velocity, angular, dims, inertia,
 * inv_inertia, mass, inv_mass, gravity_scale, damping, angular_damping, size, reach,
 * friction, restitution, hull, framed, offset, frame, field}; with `hulls`, the hull bodies'
 * hulls moved into their frames (numbered from `firstHull` in the hull column), and
 * `fields`, the heightfields ({columns, rows, spacing, heights}, numbered from `firstField`).
 */
function resolve(caller, list, firstHull, firstField) {
  const hullBase = firstHull === undefined ? 0 : firstHull, fieldBase = firstField === undefined ? 0 : firstField;
  const n = list.length;
  const out = {
    n,
    type: new Uint8Array(n), shape: new Uint8Array(n), lock: new Uint8Array(n),
    position: new Float64Array(3 * n), rotation: new Float64Array(4 * n), velocity: new Float64Array(3 * n),
    angular: new Float64Array(3 * n), dims: new Float64Array(3 * n), inertia: new Float64Array(3 * n),
    inv_inertia: new Float64Array(3 * n), mass: new Float64Array(n), inv_mass: new Float64Array(n),
    gravity_scale: new Float64Array(n), damping: new Float64Array(n), angular_damping: new Float64Array(n),
    size: new Float64Array(n), reach: new Float64Array(n), friction: new Float64Array(n),
    restitution: new Float64Array(n), hull: new Int32Array(n).fill(-1), framed: new Uint8Array(n),
    offset: new Float64Array(3 * n), frame: new Float64Array(4 * n), field: new Int32Array(n).fill(-1),
    hulls: [], fields: [],
  };
  for (let i = 0; i < n; i++) {
    const given = list[i] === undefined ? {} : list[i];
    if (given === null || typeof given !== 'object') fail(caller, 'options must be an object');
    for (const k of Object.keys(given)) {
      if (k === 'vertices' && given.shape === 'hull') fail(caller, 'a hull is taken as built (hull: {vertices, start, loops, planes}), not as points');
      if (!TAKEN.has(k)) fail(caller, `${k} is not taken by this world`);
    }
    const o = Object.assign({}, DEFAULTS, given);
    if (!(o.shape in SHAPE)) fail(caller, `a ${o.shape} is not taken by this front-end (boxes, spheres, capsules, built hulls and heightfields)`);
    const level = o.shape === 'heightfield';
    // Level geometry is static unless a type is given, and never dynamic (src/world.js:2338-2350).
    if (level && given.type === undefined) o.type = 'static';
    if (o.type !== 'dynamic' && o.type !== 'static') fail(caller, `type must be 'dynamic' or 'static', got ${o.type}`);
    if (level && o.type === 'dynamic') fail(caller, `type must be 'static' or 'kinematic' for a heightfield, got '${o.type}'`);
    if (level) for (const k of ['mass', 'density']) if (given[k] !== undefined) fail(caller, `a heightfield has no ${k}`);
    if (given.mass !== undefined && given.density !== undefined) fail(caller, 'give mass or density, not both');
    const dynamic = o.type === 'dynamic';
    let volume, inertia, size, reach, mass;
    let place = vector(caller, 'position', o.position, 3), turn = rotation(caller, o.rotation);
    if (o.shape === 'box') {
      const s = vector(caller, 'size', o.size, 3);
      const w = positive(caller, 'size', s[0]), h = positive(caller, 'size', s[1]), d = positive(caller, 'size', s[2]);
      out.dims[3 * i] = w; out.dims[3 * i + 1] = h; out.dims[3 * i + 2] = d;
      volume = w * h * d;
      mass = given.mass !== undefined ? positive(caller, 'mass', given.mass) : positive(caller, 'density', o.density) * volume;
      inertia = [mass * (h * h + d * d) / 12, mass * (w * w + d * d) / 12, mass * (w * w + h * h) / 12];
      size = Math.min(w, h, d);
      reach = 0.5 * Math.sqrt(w * w + h * h + d * d);
    } else if (o.shape === 'sphere') {
      const r = positive(caller, 'radius', o.radius);
      out.dims[3 * i] = r;
      volume = (4 / 3) * Math.PI * r * r * r;
      mass = given.mass !== undefined ? positive(caller, 'mass', given.mass) : positive(caller, 'density', o.density) * volume;
      inertia = [0.4 * mass * r * r, 0.4 * mass * r * r, 0.4 * mass * r * r];
      size = 2 * r;
      reach = r;
    } else if (o.shape === 'capsule') {
      // src/shapes.js:14-21 (volume) and 37-55 (inertia): a cylinder of length height - 2r
      // along y, capped by two hemispheres, the mass shared between them by volume.
      const r = positive(caller, 'radius', o.radius), height = positive(caller, 'height', o.height);
      if (height < 2 * r) fail(caller, `height must be at least twice radius (${2 * r}), got ${height}`);
      out.dims[3 * i] = r; out.dims[3 * i + 1] = height;
      const ball = (4 / 3) * Math.PI * r * r * r;
      volume = Math.PI * r * r * (height - 2 * r) + ball;
      mass = given.mass !== undefined ? positive(caller, 'mass', given.mass) : positive(caller, 'density', o.density) * volume;
      const L = height - 2 * r, cylinder = Math.PI * r * r * L;
      const mc = mass * cylinder / (cylinder + ball), ms = mass * ball / (cylinder + ball);
      const across = mc * (L * L / 12 + r * r / 4) + ms * (0.4 * r * r + L * L / 4 + 3 * L * r / 8);
      inertia = [across, mc * r * r / 2 + ms * 0.4 * r * r, across];
      size = 2 * r;
      reach = height * 0.5;
    } else if (o.shape === 'heightfield') {
      // src/mesh.js:288-305 (Heightfield): centred on the body, x along columns, z along rows.
      const columns = whole(caller, 'columns', o.columns, 2);
      const spacing = positive(caller, 'spacing', o.spacing);
      const heights = numbers(caller, 'heights', o.heights, Float64Array);
      if (heights.length % columns !== 0 || heights.length / columns < 2) {
        fail(caller, `heights must be whole rows of ${columns}, at least 2 of them`);
      }
      const rows = heights.length / columns;
      const x0 = -((columns - 1) * spacing) / 2, z0 = -((rows - 1) * spacing) / 2;
      let lo = Infinity, hi = -Infinity;
      for (const h of heights) { lo = Math.min(lo, h); hi = Math.max(hi, h); }
      out.field[i] = fieldBase + out.fields.length;
      out.fields.push({ columns, rows, spacing, heights });
      volume = 0;
      mass = positive(caller, 'density', o.density) * volume;
      inertia = [0, 0, 0];
      size = 0;
      reach = Math.sqrt(x0 * x0 + z0 * z0 + Math.max(lo * lo, hi * hi));
    } else {
      // A hull (src/world.js:2167-2178 and 2544-2549 solidMass): its mass properties for a
      // density of 1 scaled to the body's, the principal axes its frame, its centre of mass
      // the offset along them (src/world.js:2194-2202), the hull moved there.
      const built = builtHull(caller, o.hull);
      const hm = hullMass(built);
      volume = hm.volume;
      mass = given.mass !== undefined ? positive(caller, 'mass', given.mass) : positive(caller, 'density', o.density) * hm.volume;
      const density = mass / hm.volume;
      const axes = principal(hm.tensor.map((x) => x * density));
      inertia = axes.moments;
      const R = matrixOf(axes.rotation), c = hm.center;
      const inner = moved(built, c, transposed(R));
      out.hull[i] = hullBase + out.hulls.length;
      out.hulls.push(inner);
      out.framed[i] = 1;
      out.frame.set(axes.rotation, 4 * i);
      const offset = [0, 1, 2].map((k) => R[k] * c[0] + R[3 + k] * c[1] + R[6 + k] * c[2]);
      out.offset.set(offset, 3 * i);
      size = inner.size;
      reach = inner.reach;
      // src/world.js:1933-1945 (_place7): the step's pose is the centre of mass, turned to the axes.
      const q = quatMultiply(turn, axes.rotation);
      const d = rotate(q, offset[0], offset[1], offset[2]);
      place = [place[0] + d[0], place[1] + d[1], place[2] + d[2]];
      turn = q;
    }
    out.type[i] = dynamic ? 0 : 1;
    out.shape[i] = SHAPE[o.shape];
    if (!out.framed[i]) out.frame[4 * i + 3] = 1;
    out.position.set(place, 3 * i);
    out.rotation.set(turn, 4 * i);
    out.velocity.set(vector(caller, 'velocity', o.velocity, 3), 3 * i);
    out.angular.set(vector(caller, 'angularVelocity', o.angularVelocity, 3), 3 * i);
    for (let a = 0; a < 3; a++) {
      out.inertia[3 * i + a] = inertia[a];
      out.inv_inertia[3 * i + a] = inertia[a] > 0 ? 1 / inertia[a] : 0;
    }
    out.mass[i] = mass;
    out.inv_mass[i] = dynamic ? 1 / mass : 0;
    out.gravity_scale[i] = o.gravityScale;
    out.damping[i] = o.damping;
    out.angular_damping[i] = o.angularDamping;
    out.size[i] = size;
    out.reach[i] = reach;
    out.friction[i] = o.friction;
    out.restitution[i] = o.restitution;
  }
  return out;
}

module.exports = { DEFAULTS, resolve };
