// The hulls of a scene's hull bodies as the reference builds them (src/hull.js build(), in
// the frame their points are given in, before a world moves them to their centre of mass):
// for the JavaScript front-end's tests, which take hulls as built tables. Per hull body, in
// the scene's order: its body number, its corners, faces (start, one more than the faces),
// loops and planes, joined with offset tables (V corners, F faces, L loop entries, n + 1
// each). Run through lower.mjs like make_fixtures.mjs.
import { build } from '/root/reference/src/hull.js';
import { readFileSync, writeFileSync } from 'fs';
import { gzipSync } from 'zlib';

const scene = JSON.parse(readFileSync(process.argv[2], 'utf8'));
const body = [], vertices = [], start = [], loops = [], planes = [], V = [0], F = [0], L = [0];
scene.forEach((spec, i) => {
  if (spec.shape !== 'hull') return;
  const h = build(spec.vertices);
  if (h === null) throw new Error(`make_built: body ${i} spans no solid`);
  body.push(i);
  for (const x of h.vertices) vertices.push(x);
  for (const x of h.start) start.push(x);
  for (const x of h.loops) loops.push(x);
  for (const x of h.planes) planes.push(x);
  V.push(vertices.length / 3); F.push(F[F.length - 1] + h.start.length - 1); L.push(loops.length);
});
const arrays = {
  body: Int32Array.from(body), vertices: Float64Array.from(vertices), start: Int32Array.from(start),
  loops: Int32Array.from(loops), planes: Float64Array.from(planes), V: Int32Array.from(V), F: Int32Array.from(F),
  L: Int32Array.from(L),
};
const entries = [];
let offset = 0;
for (const [name, v] of Object.entries(arrays)) {
  entries.push({ name, dtype: v instanceof Float64Array ? 'f64' : 'i32', length: v.length, offset });
  offset += Math.ceil(v.byteLength / 8) * 8;
}
const header = Buffer.from(JSON.stringify({ arrays: entries, meta: { hulls: body.length } }));
const pad = (8 - ((12 + header.length) % 8)) % 8;
const out = Buffer.alloc(12 + header.length + pad + offset);
out.write('BMFX0001', 0, 'latin1'); out.writeUInt32LE(header.length, 8); header.copy(out, 12);
let k = 0;
for (const v of Object.values(arrays)) Buffer.from(v.buffer, v.byteOffset, v.byteLength).copy(out, 12 + header.length + pad + entries[k++].offset);
writeFileSync(process.argv[3], gzipSync(out, { level: 9 }));
