'use strict';
/*
 * Drives the JavaScript front-end (bearing_amd/js) for the Python tests, one mode a run:
 *   resolve <scene.json>                 the scene's addBody columns as JSON (no addon, no GPU)
 *   load                                 loads the addon, prints its exports and what new World() threw, if anything
 *   steps <scene.json> <steps> <dt>      steps the scene: moved(0)'s digest, then one JSON line a step:
 *                                        [sha256 of transforms, awake, contacts, frozen, of moved() twice, of previous]
 */
const fs = require('fs');
const path = require('path');
const crypto = require('crypto');
const front = path.join(__dirname, '..', '..', 'bearing_amd', 'js');

const [mode, scene, steps, dt] = process.argv.slice(2);
if (mode === 'resolve') {
  const { resolve } = require(path.join(front, 'resolve.js'));
  const cols = resolve('addBody', JSON.parse(fs.readFileSync(scene, 'utf8')));
  const out = {};
  for (const k of Object.keys(cols)) out[k] = k === 'n' ? cols.n : Array.from(cols[k]);
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
  // moved(stepCount - 7, or 0), of previous]; each moved() list hashed as its Int32Array, in order.
  const { World } = require(path.join(front, 'world.js'));
  const world = new World();
  world.addBodies(JSON.parse(fs.readFileSync(scene, 'utf8')));
  const digest = (a) => crypto.createHash('sha256').update(Buffer.from(a.buffer, a.byteOffset, a.byteLength)).digest('hex');
  process.stdout.write(JSON.stringify([digest(world.moved(0))]) + '\n');
  const n = Number(steps), h = Number(dt);
  for (let i = 0; i < n; i++) {
    world.step(h);
    const c = world.counts;
    const last = digest(world.moved(world.stepCount - 1));
    const older = digest(world.moved(Math.max(0, world.stepCount - 7)));
    process.stdout.write(JSON.stringify([digest(world.transforms), c.awake, c.contacts, c.frozen, last, older,
      digest(world.previous)]) + '\n');
  }
  world.destroy();
} else {
  throw new Error(`run_scene: unknown mode ${mode}`);
}
