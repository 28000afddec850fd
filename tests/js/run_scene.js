'use strict';
/*
 * Drives the JavaScript front-end (bearing_amd/js) for the Python tests, one mode a run:
 *   resolve <scene.json> [built]         the scene's addBody columns as JSON, with its hull and heightfield tables
 *                                        joined as the world joins them (no addon, no GPU)
 *   load                                 loads the addon, prints its exports and what new World() threw, if anything
 *   steps <scene.json> <steps> <dt> [built]  steps the scene: moved(0)'s digest, then one JSON line a step:
 *                                        [sha256 of transforms, awake, contacts, frozen, of moved() twice, of previous,
 *                                        of the step's contact events]
 * `built`: a fixture of the scene's hulls as the reference's src/hull.js build() made them
 * (tests/golden/make_built.mjs); each hull body's points are replaced by its built hull.
 */
const fs = require('fs');
const path = require('path');
const crypto = require('crypto');
const front = path.join(__dirname, '..', '..', 'bearing_amd', 'js');

const zlib = require('zlib');

/* A BMFX fixture (tests/golden/fixtures.py): its arrays by name. */
function fixture(file) {
  const raw = zlib.gunzipSync(fs.readFileSync(file));
  if (raw.toString('latin1', 0, 8) !== 'BMFX0001') throw new Error(`${file}: not a BMFX fixture`);
  const hlen = raw.readUInt32LE(8);
  const header = JSON.parse(raw.toString('utf8', 12, 12 + hlen));
  const base = 12 + hlen + ((8 - ((12 + hlen) % 8)) % 8);
  const kinds = { f64: Float64Array, i32: Int32Array, u32: Uint32Array, u8: Uint8Array };
  const out = {};
  for (const e of header.arrays) {
    const K = kinds[e.dtype];
    const bytes = raw.slice(base + e.offset, base + e.offset + e.length * K.BYTES_PER_ELEMENT);
    out[e.name] = new K(Uint8Array.from(bytes).buffer);
  }
  return out;
}

/* The scene's bodies, each hull body's points replaced by its built hull from `built`. */
function bodies(sceneFile, builtFile) {
  const list = JSON.parse(fs.readFileSync(sceneFile, 'utf8'));
  if (!builtFile) return list;
  const b = fixture(builtFile);
  for (let k = 0; k < b.body.length; k++) {
    const spec = Object.assign({}, list[b.body[k]]);
    delete spec.vertices;
    spec.hull = {
      vertices: b.vertices.subarray(3 * b.V[k], 3 * b.V[k + 1]),
      start: b.start.subarray(b.F[k] + k, b.F[k + 1] + k + 1),
      loops: b.loops.subarray(b.L[k], b.L[k + 1]),
      planes: b.planes.subarray(4 * b.F[k], 4 * b.F[k + 1]),
    };
    list[b.body[k]] = spec;
  }
  return list;
}

const [mode, scene, steps, dt] = process.argv.slice(2);
if (mode === 'resolve') {
  const { resolve } = require(path.join(front, 'resolve.js'));
  const { joinHulls, joinFields } = require(path.join(front, 'world.js'));
  const cols = resolve('addBody', bodies(scene, steps));
  const out = {};
  for (const k of Object.keys(cols)) {
    if (k === 'hulls' || k === 'fields') continue;
    out[k] = k === 'n' ? cols.n : Array.from(cols[k]);
  }
  const tables = { hulls: joinHulls(cols.hulls), fields: joinFields(cols.fields) };
  for (const [name, t] of Object.entries(tables)) {
    out[name] = {};
    for (const k of Object.keys(t)) out[name][k] = Array.from(t[k]);
  }
  process.stdout.write(JSON.stringify(out));
} else if (mode === 'load') {
  const addon = require(path.join(front, '..', '_build', 'bearing_napi.node'));
  let threw = null;
  try {
    const { World } = require(path.join(front, 'world.js'));
    new World().destroy();
  } catch (e) {
    threw = e.message;
  }
  process.stdout.write(JSON.stringify({ exports: Object.keys(addon), threw }));
} else if (mode === 'steps') {
  // One JSON line before the first step, the SHA-256 of moved(0) (every body added), then one a
  // step: [sha256 of transforms, awake, contacts, frozen, sha256 of moved(stepCount - 1), of
  // moved(stepCount - 7, or 0), of previous, of the contact events]; each moved() list hashed as
  // its Int32Array, in order.
  const { World } = require(path.join(front, 'world.js'));
  const world = new World();
  world.addBodies(bodies(scene, process.argv[6]));
  const digest = (a) => crypto.createHash('sha256').update(Buffer.from(a.buffer, a.byteOffset, a.byteLength)).digest('hex');
  process.stdout.write(JSON.stringify([digest(world.moved(0))]) + '\n');
  const n = Number(steps), h = Number(dt);
  for (let i = 0; i < n; i++) {
    world.step(h);
    const c = world.counts;
    const last = digest(world.moved(world.stepCount - 1));
    const older = digest(world.moved(Math.max(0, world.stepCount - 7)));
    // The step's contact events as the reference's fixture hashes them: ten numbers an event.
    const ev = world.contacts(), rows = new Float64Array(10 * ev.length);
    ev.forEach((x, j) => {
      rows.set([x.type === 'end' ? 1 : 0, x.a, x.b], 10 * j);
      if (x.type === 'begin') rows.set([...x.point, ...x.normal, x.speed], 10 * j + 3);
    });
    process.stdout.write(JSON.stringify([digest(world.transforms), c.awake, c.contacts, c.frozen, last, older,
      digest(world.previous), digest(rows)]) + '\n');
  }
  world.destroy();
} else {
  throw new Error(`run_scene: unknown mode ${mode}`);
}
