// The reference's touchLevel (src/triangle.js:1078) on seeded shapes against heightfields
// (src/mesh.js Heightfield), written by make_fixtures.py (level command, JSON), and every
// field triangle's edge flags and folds (Heightfield.edge). Hulls are built by the
// reference's own src/hull.js build(); their tables are written beside the manifolds. Run
// through lower.mjs like make_fixtures.mjs.
import { touchLevel, Touches } from '/root/reference/src/triangle.js';
import { Heightfield } from '/root/reference/src/mesh.js';
import { Shape, HULL } from '/root/reference/src/collide.js';
import { build } from '/root/reference/src/hull.js';
import { PIXEL } from '/root/reference/src/contact.js';
import { readFileSync, writeFileSync } from 'fs';
import { gzipSync } from 'zlib';

const job = JSON.parse(readFileSync(process.argv[2], 'utf8'));
const CAP = 32;
const fields = job.fields.map((f) => new Heightfield(f.heights, f.columns, f.spacing));
const tables = { vertices: [], start: [], loops: [], planes: [], hullV: [0], hullF: [0], hullL: [0] };
let hulls = 0;
const addHull = (h) => {
  for (const x of h.vertices) tables.vertices.push(x);
  for (let f = 0; f < h.start.length - 1; f++) tables.start.push(h.start[f]);
  for (const x of h.loops) tables.loops.push(x);
  for (const x of h.planes) tables.planes.push(x);
  tables.hullV.push(tables.vertices.length / 3);
  tables.hullF.push(tables.planes.length / 4);
  tables.hullL.push(tables.loops.length);
  return hulls++;
};
const n = job.cases.length;
const shapes = new Float64Array(19 * n), boxes = new Float64Array(6 * n), margin = new Float64Array(n);
const tolerance = new Float64Array(n), field = new Int32Array(n), position = new Float64Array(3 * n);
const axes = new Float64Array(9 * n), manifolds = new Int32Array(n), count = new Int32Array(CAP * n);
const normal = new Float64Array(3 * CAP * n), points = new Float64Array(12 * CAP * n);
const separations = new Float64Array(4 * CAP * n), ids = new Uint32Array(4 * CAP * n);
const touches = new Touches();
const s = new Shape();
for (let k = 0; k < n; k++) {
  const c = job.cases[k], spec = c.shape;
  s.kind = spec.kind;
  s.hull = null;
  s.half[0] = 0; s.half[1] = 0; s.half[2] = 0;
  let size;
  if (spec.kind === HULL) {
    s.hull = build(spec.points);
    const id = addHull(s.hull);
    size = s.hull.size;
    shapes[19 * k + 13] = id;
  } else {
    s.half.set(spec.half);
    size = spec.kind === 2 ? 2 * Math.min(spec.half[0], spec.half[1], spec.half[2]) : 2 * spec.half[0];
    shapes.set(spec.half, 19 * k + 13);
  }
  s.place(spec.center[0], spec.center[1], spec.center[2], spec.q, 0);
  shapes[19 * k] = s.kind; shapes.set(s.center, 19 * k + 1); shapes.set(s.axes, 19 * k + 4);
  field[k] = c.field; position.set(c.position, 3 * k); axes.set(c.axes, 9 * k); boxes.set(c.box, 6 * k);
  margin[k] = c.margin;
  tolerance[k] = size * PIXEL;
  const got = touchLevel(fields[c.field], Float64Array.from(c.position), Float64Array.from(c.axes), 0, s,
    c.box[0], c.box[1], c.box[2], c.box[3], c.box[4], c.box[5], margin[k], tolerance[k], touches);
  if (got > CAP) throw new Error(`case ${k}: ${got} manifolds`);
  manifolds[k] = got;
  for (let j = 0; j < got; j++) {
    const m = touches.list[j], o = CAP * k + j;
    count[o] = m.count;
    normal.set(m.normal, 3 * o);
    for (let i = 0; i < m.count; i++) {
      points.set(m.points.subarray(3 * i, 3 * i + 3), 12 * o + 3 * i);
      separations[4 * o + i] = m.separations[i];
      ids[4 * o + i] = m.ids[i];
    }
  }
}
const arrays = { shapes, boxes, margin, tolerance, field, position, axes, manifolds, count, normal, points, separations, ids };
job.fields.forEach((f, i) => {
  const g = fields[i], T = g.count, flags = new Int32Array(T), folds = new Float64Array(6 * T), fo = new Float64Array(6);
  for (let t = 0; t < T; t++) { flags[t] = g.edge(t, fo); folds.set(fo, 6 * t); }
  arrays[`field${i}.heights`] = Float64Array.from(f.heights);
  arrays[`field${i}.flags`] = flags;
  arrays[`field${i}.folds`] = folds;
});
Object.assign(arrays, {
  vertices: Float64Array.from(tables.vertices), start: Int32Array.from(tables.start), loops: Int32Array.from(tables.loops),
  planes: Float64Array.from(tables.planes), hullV: Int32Array.from(tables.hullV), hullF: Int32Array.from(tables.hullF),
  hullL: Int32Array.from(tables.hullL),
});
const entries = [];
let offset = 0;
for (const [name, v] of Object.entries(arrays)) {
  entries.push({ name, dtype: v instanceof Float64Array ? 'f64' : v instanceof Int32Array ? 'i32' : 'u32', length: v.length, offset });
  offset += Math.ceil(v.byteLength / 8) * 8;
}
const meta = { cases: n, cap: CAP, hulls, fields: job.fields.map((f) => ({ columns: f.columns, rows: f.rows, spacing: f.spacing })) };
const header = Buffer.from(JSON.stringify({ arrays: entries, meta }));
const pad = (8 - ((12 + header.length) % 8)) % 8;
const out = Buffer.alloc(12 + header.length + pad + offset);
out.write('BMFX0001', 0, 'latin1'); out.writeUInt32LE(header.length, 8); header.copy(out, 12);
let k = 0;
for (const v of Object.values(arrays)) Buffer.from(v.buffer, v.byteOffset, v.byteLength).copy(out, 12 + header.length + pad + entries[k++].offset);
writeFileSync(process.argv[3], gzipSync(out, { level: 9 }));
