// The reference's time of impact (src/impact.js impact), core distance (src/gjk.js
// closest) and trigonometry (src/core/math.js sin, cos) on seeded pairs written by
// make_fixtures.py (impact command), for the oracle's fixtures. Each pair's two bodies
// sit in slots 0 and 1 of a world-shaped object holding just the columns impact() reads.
// Run through lower.mjs like make_fixtures.mjs.
import { impact } from '/root/reference/src/impact.js';
import { closest, Closest } from '/root/reference/src/gjk.js';
import { Shape } from '/root/reference/src/collide.js';
import { sin, cos } from '/root/reference/src/core/math.js';
import { readFileSync, writeFileSync } from 'fs';
import { gzipSync } from 'zlib';

const PER = 20, REC = 2 * PER + 5;
const raw = readFileSync(process.argv[2]);
const input = new Float64Array(raw.buffer, raw.byteOffset, raw.byteLength / 8);
const n = input.length / REC;
const bodies = new Float64Array(n * 2 * 32), dt = new Float64Array(n), tolerance = new Float64Array(n);
const gravity = new Float64Array(3 * n);
const distance = new Float64Array(n), nearA = new Float64Array(3 * n), nearB = new Float64Array(3 * n);
const t = new Float64Array(n), posedA = new Float64Array(12 * n), posedB = new Float64Array(12 * n);
const world = {
  _position: new Float64Array(6), _velocity: new Float64Array(6), _angular: new Float64Array(6),
  _rotation: new Float64Array(8), _type: new Uint8Array(2), _gravityScale: new Float64Array(2),
  _gravity: new Float64Array(3), _shape: new Uint8Array(2), _framed: new Uint8Array(2), _reach: new Float64Array(2),
};
const sa = new Shape(), sb = new Shape(), near = new Closest();
for (let k = 0; k < n; k++) {
  const r = input.subarray(REC * k, REC * (k + 1));
  dt[k] = r[2 * PER]; tolerance[k] = r[2 * PER + 1];
  world._gravity.set(r.subarray(2 * PER + 2, 2 * PER + 5));
  gravity.set(world._gravity, 3 * k);
  [sa, sb].forEach((s, i) => {
    const o = PER * i, kind = r[o], dynamic = r[o + 18] === 1, scale = dynamic ? r[o + 17] : 0;
    world._position.set(r.subarray(o + 1, o + 4), 3 * i);
    world._rotation.set(r.subarray(o + 4, o + 8), 4 * i);
    world._velocity.set(r.subarray(o + 11, o + 14), 3 * i);
    world._angular.set(r.subarray(o + 14, o + 17), 3 * i);
    world._gravityScale[i] = r[o + 17];
    world._type[i] = dynamic ? 0 : 1;
    world._shape[i] = kind === 0 ? 1 : 0;
    world._reach[i] = r[o + 19];
    s.kind = kind;
    s.half.set(r.subarray(o + 8, o + 11));
    s.place(r[o + 1], r[o + 2], r[o + 3], world._rotation, 4 * i);
    for (let c = 0; c < 3; c++) s.weight[c] = world._gravity[c] * scale;
    // The oracle's body record: shape (kind, centre, axes, half, weight), q, v, w,
    // gravity scale as moving() reads it, reach, flags (1 dynamic, 2 spin capped).
    const d = 32 * (2 * k + i);
    bodies[d] = kind; bodies.set(s.center, d + 1); bodies.set(s.axes, d + 4); bodies.set(s.half, d + 13);
    bodies.set(s.weight, d + 16); bodies.set(r.subarray(o + 4, o + 8), d + 19); bodies.set(r.subarray(o + 11, o + 17), d + 23);
    bodies[d + 29] = scale; bodies[d + 30] = r[o + 19];
    bodies[d + 31] = (dynamic ? 1 : 0) | (dynamic && kind !== 0 ? 2 : 0);
  });
  distance[k] = closest(sa, sb, near);
  nearA.set(near.a, 3 * k); nearB.set(near.b, 3 * k);
  t[k] = impact(world, 0, 1, sa, sb, dt[k], tolerance[k], null, null);
  posedA.set(sa.center, 12 * k); posedA.set(sa.axes, 12 * k + 3);
  posedB.set(sb.center, 12 * k); posedB.set(sb.axes, 12 * k + 3);
}
// Trigonometry: a spread of arguments, small ones, near multiples of pi/2, and large ones.
const xs = [];
let seed = 2026;
const next = () => { seed = (seed * 1103515245 + 12345) % 2147483648; return seed / 2147483648; };
for (let i = 0; i < 2000; i++) xs.push((next() - 0.5) * 20);
for (let i = 0; i < 500; i++) xs.push((next() - 0.5) * 1e-3);
for (let i = 1; i < 400; i++) xs.push(i * Math.PI / 2, i * Math.PI / 2 * (1 + 1e-15), -i * Math.PI / 2);
for (let i = 0; i < 500; i++) xs.push((next() - 0.5) * 2e5);
xs.push(0, -0, 1e-300, Math.PI / 4, -Math.PI / 4, 0.5, -0.5);
const x = Float64Array.from(xs), sine = x.map((v) => sin(v)), cosine = x.map((v) => cos(v));
const arrays = { bodies, dt, tolerance, gravity, distance, nearA, nearB, t, posedA, posedB, x, sine, cosine };
const entries = [];
let offset = 0;
for (const [name, v] of Object.entries(arrays)) {
  entries.push({ name, dtype: 'f64', length: v.length, offset });
  offset += v.byteLength;
}
const header = Buffer.from(JSON.stringify({ arrays: entries, meta: { pairs: n } }));
const pad = (8 - ((12 + header.length) % 8)) % 8;
const out = Buffer.alloc(12 + header.length + pad + offset);
out.write('BMFX0001', 0, 'latin1'); out.writeUInt32LE(header.length, 8); header.copy(out, 12);
let k = 0;
for (const v of Object.values(arrays)) Buffer.from(v.buffer, v.byteOffset, v.byteLength).copy(out, 12 + header.length + pad + entries[k++].offset);
writeFileSync(process.argv[3], gzipSync(out, { level: 9 }));
