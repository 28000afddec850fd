// The reference's collide() (src/collide.js) on seeded shape pairs, for the narrowphase
// fixtures. Reads pairs written by make_fixtures.py (Float64 records: per shape kind,
// centre 3, rotation 4, half 3, weight 3; then margin, tolerance), places each shape as
// the reference places it (Shape.place) and writes the placed axes and the manifolds.
// Run through lower.mjs like make_fixtures.mjs.
import { collide, Shape, Manifold } from '/root/reference/src/collide.js';
import { readFileSync, writeFileSync } from 'fs';
import { gzipSync } from 'zlib';

const REC = 2 * 14 + 2;
const raw = readFileSync(process.argv[2]);
const input = new Float64Array(raw.buffer, raw.byteOffset, raw.byteLength / 8);
const n = input.length / REC;
const shapes = new Float64Array(n * 2 * 19), margin = new Float64Array(n), tolerance = new Float64Array(n);
const count = new Int32Array(n), normal = new Float64Array(3 * n), points = new Float64Array(12 * n);
const separations = new Float64Array(4 * n), ids = new Uint32Array(4 * n);
const a = new Shape(), b = new Shape(), m = new Manifold();
const q = new Float64Array(4);
for (let k = 0; k < n; k++) {
  const r = input.subarray(REC * k, REC * (k + 1));
  [a, b].forEach((s, t) => {
    const o = 14 * t;
    s.kind = r[o];
    q.set(r.subarray(o + 4, o + 8));
    s.place(r[o + 1], r[o + 2], r[o + 3], q, 0);
    s.half.set(r.subarray(o + 8, o + 11));
    s.weight.set(r.subarray(o + 11, o + 14));
    const d = 19 * (2 * k + t);
    shapes[d] = s.kind; shapes.set(s.center, d + 1); shapes.set(s.axes, d + 4); shapes.set(s.half, d + 13);
    shapes.set(s.weight, d + 16);
  });
  margin[k] = r[28]; tolerance[k] = r[29];
  m.count = 0;
  count[k] = collide(a, b, margin[k], tolerance[k], m);
  normal.set(m.normal, 3 * k); points.set(m.points, 12 * k); separations.set(m.separations, 4 * k); ids.set(m.ids, 4 * k);
}
const arrays = { shapes, margin, tolerance, count, normal, points, separations, ids };
const entries = [];
let offset = 0;
for (const [name, v] of Object.entries(arrays)) {
  entries.push({ name, dtype: v instanceof Float64Array ? 'f64' : v instanceof Int32Array ? 'i32' : 'u32', length: v.length, offset });
  offset += Math.ceil(v.byteLength / 8) * 8;
}
const header = Buffer.from(JSON.stringify({ arrays: entries, meta: { pairs: n } }));
const pad = (8 - ((12 + header.length) % 8)) % 8;
const out = Buffer.alloc(12 + header.length + pad + offset);
out.write('BMFX0001', 0, 'latin1'); out.writeUInt32LE(header.length, 8); header.copy(out, 12);
let k = 0;
for (const v of Object.values(arrays)) Buffer.from(v.buffer, v.byteOffset, v.byteLength).copy(out, 12 + header.length + pad + entries[k++].offset);
writeFileSync(process.argv[3], gzipSync(out, { level: 9 }));
