// The reference's narrowphase with hulls (src/collide.js hullHull, sphereHull, capsuleHull,
// and a box against a hull, which is placed as src/collide.js's CUBE) on seeded pairs
// written by make_fixtures.py (hulls command, JSON). Each hull is built from its points by
// the reference's own src/hull.js build(); the built tables are written beside the
// manifolds, with the unit cube's as build() makes it. Run through lower.mjs like
// make_fixtures.mjs.
import { collide, Shape, Manifold, HULL } from '/root/reference/src/collide.js';
import { build } from '/root/reference/src/hull.js';
import { PIXEL } from '/root/reference/src/contact.js';
import { readFileSync, writeFileSync } from 'fs';
import { gzipSync } from 'zlib';

const pairs = JSON.parse(readFileSync(process.argv[2], 'utf8'));
const hulls = [];
const tables = { vertices: [], start: [], loops: [], planes: [], hullV: [0], hullF: [0], hullL: [0] };
const addHull = (h) => {
  hulls.push(h);
  for (const x of h.vertices) tables.vertices.push(x);
  for (let f = 0; f < h.start.length - 1; f++) tables.start.push(h.start[f]);
  for (const x of h.loops) tables.loops.push(x);
  for (const x of h.planes) tables.planes.push(x);
  tables.hullV.push(tables.vertices.length / 3);
  tables.hullF.push(tables.planes.length / 4);
  tables.hullL.push(tables.loops.length);
  return hulls.length - 1;
};
addHull(build([-0.5, -0.5, -0.5, 0.5, -0.5, -0.5, -0.5, 0.5, -0.5, 0.5, 0.5, -0.5,
  -0.5, -0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5, 0.5, 0.5, 0.5]));   // src/collide.js CUBE
const n = pairs.length;
const shapes = new Float64Array(n * 2 * 19), margin = new Float64Array(n), tolerance = new Float64Array(n);
const count = new Int32Array(n), normal = new Float64Array(3 * n), points = new Float64Array(12 * n);
const separations = new Float64Array(4 * n), ids = new Uint32Array(4 * n), loss = new Float64Array(n);
const m = new Manifold();
const made = [new Shape(), new Shape()];
for (let k = 0; k < n; k++) {
  const pr = pairs[k];
  const size = [];
  [pr.a, pr.b].forEach((spec, t) => {
    const s = made[t];
    s.kind = spec.kind;
    s.hull = null;
    s.placed = null;
    s.half[0] = 0; s.half[1] = 0; s.half[2] = 0;
    let id = -1;
    if (spec.kind === HULL) {
      const h = build(spec.points);
      s.hull = h;
      id = addHull(h);
      size.push(h.size);
    } else {
      s.half.set(spec.half);
      size.push(spec.kind === 2 ? 2 * Math.min(spec.half[0], spec.half[1], spec.half[2]) : 2 * spec.half[0]);
    }
    s.place(spec.center[0], spec.center[1], spec.center[2], spec.q, 0);
    s.weight.set(spec.weight);
    const d = 19 * (2 * k + t);
    shapes[d] = s.kind; shapes.set(s.center, d + 1); shapes.set(s.axes, d + 4);
    if (spec.kind === HULL) shapes[d + 13] = id; else shapes.set(s.half, d + 13);
    shapes.set(s.weight, d + 16);
  });
  margin[k] = pr.margin;
  tolerance[k] = Math.min(size[0], size[1]) * PIXEL;
  m.count = 0;
  count[k] = collide(made[0], made[1], margin[k], tolerance[k], m);
  normal.set(m.normal, 3 * k); points.set(m.points, 12 * k); separations.set(m.separations, 4 * k); ids.set(m.ids, 4 * k);
  loss[k] = m.loss;
}
const arrays = {
  shapes, margin, tolerance, count, normal, points, separations, ids, loss,
  vertices: Float64Array.from(tables.vertices), start: Int32Array.from(tables.start), loops: Int32Array.from(tables.loops),
  planes: Float64Array.from(tables.planes), hullV: Int32Array.from(tables.hullV), hullF: Int32Array.from(tables.hullF),
  hullL: Int32Array.from(tables.hullL),
};
const entries = [];
let offset = 0;
for (const [name, v] of Object.entries(arrays)) {
  entries.push({ name, dtype: v instanceof Float64Array ? 'f64' : v instanceof Int32Array ? 'i32' : 'u32', length: v.length, offset });
  offset += Math.ceil(v.byteLength / 8) * 8;
}
const header = Buffer.from(JSON.stringify({ arrays: entries, meta: { pairs: n, hulls: hulls.length } }));
const pad = (8 - ((12 + header.length) % 8)) % 8;
const out = Buffer.alloc(12 + header.length + pad + offset);
out.write('BMFX0001', 0, 'latin1'); out.writeUInt32LE(header.length, 8); header.copy(out, 12);
let k = 0;
for (const v of Object.values(arrays)) Buffer.from(v.buffer, v.byteOffset, v.byteLength).copy(out, 12 + header.length + pad + entries[k++].offset);
writeFileSync(process.argv[3], gzipSync(out, { level: 9 }));
