// Writes the contact events of a trajectory by running the reference (bearing-physics, mounted
// read-only at /root/reference) on this image's Node 12 through lower.mjs. Run from the
// repository root:
//   python tests/golden/make_fixtures.py events [name ...]
// which calls
//   node --experimental-loader ./tests/golden/lower.mjs tests/golden/make_events.mjs <job.json>
// with the job make_fixtures.py wrote for the trajectory (scene, dt, steps, out).
//
// What is recorded, into <out>.events.bmfx.gz: after every step, world._events as the step
// left it (src/events.js; cleared as a step begins, src/world.js:888): the count, and each
// event as ten numbers [type, slot a, slot b, point 3, normal 3, speed] (type 0 begin, 1 end;
// an end's point, normal and speed as 0, since the reference leaves whatever its columns held
// there), with each step's events hashed (SHA-256 of its Float64Array). Every event is kept
// when there are at most KEEP of them over the whole trajectory; past that, only the hashes.
import { World } from '/root/reference/src/bearing.js';
import { createHash } from 'crypto';
import { readFileSync, writeFileSync } from 'fs';
import { gzipSync } from 'zlib';

const KEEP = 60000;

function pack(arrays, meta) {
  const entries = [];
  let offset = 0;
  for (const [name, a] of Object.entries(arrays)) {
    const dtype = a instanceof Float64Array ? 'f64' : 'i32';
    entries.push({ name, dtype, length: a.length, offset });
    offset += Math.ceil(a.byteLength / 8) * 8;
  }
  const header = Buffer.from(JSON.stringify({ arrays: entries, meta }));
  const pad = (8 - ((12 + header.length) % 8)) % 8;
  const out = Buffer.alloc(12 + header.length + pad + offset);
  out.write('BMFX0001', 0, 'latin1');
  out.writeUInt32LE(header.length, 8);
  header.copy(out, 12);
  const base = 12 + header.length + pad;
  let k = 0;
  for (const a of Object.values(arrays)) Buffer.from(a.buffer, a.byteOffset, a.byteLength).copy(out, base + entries[k++].offset);
  return gzipSync(out, { level: 9 });
}

function main() {
  const job = JSON.parse(readFileSync(process.argv[2], 'utf8'));
  const scene = JSON.parse(readFileSync(job.scene, 'utf8'));
  const world = new World({ workers: 0 });
  for (const spec of scene) world.addBody(spec);
  const counts = new Int32Array(job.steps), hashes = [], rows = [];
  let total = 0;
  for (let s = 1; s <= job.steps; s++) {
    world.step(job.dt);
    const e = world._events, n = e.count, out = new Float64Array(10 * n);
    for (let i = 0; i < n; i++) {
      const r = 10 * i, i3 = 3 * i;
      out[r] = e.type[i];
      // id = generation * 2^24 + slot (src/core/slots.js): the slot is the low 24 bits.
      out[r + 1] = e.idA[i] % 16777216;
      out[r + 2] = e.idB[i] % 16777216;
      if (e.type[i] === 0) {
        for (let k = 0; k < 3; k++) { out[r + 3 + k] = e.point[i3 + k]; out[r + 6 + k] = e.normal[i3 + k]; }
        out[r + 9] = e.speed[i];
      }
    }
    counts[s - 1] = n;
    total += n;
    hashes.push(createHash('sha256').update(Buffer.from(out.buffer, out.byteOffset, out.byteLength)).digest('hex'));
    if (total <= KEEP) rows.push(out);
  }
  const arrays = { counts };
  if (total <= KEEP) {
    const all = new Float64Array(10 * total);
    let at = 0;
    for (const r of rows) { all.set(r, at); at += r.length; }
    arrays.events = all;
  }
  writeFileSync(`${job.out}.events.bmfx.gz`, pack(arrays, { steps: job.steps, dt: job.dt, total, hashes }));
}

main();
