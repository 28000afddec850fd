'use strict';
/*
 * Drives the JavaScript front-end (bearing_amd/js) for the Python tests, one mode a run:
 *   resolve <scene.json>                 the scene's addBody columns as JSON (no addon, no GPU)
 *   load                                 loads the addon, prints its exports and what new World() threw, if anything
 *   steps <scene.json> <steps> <dt>      steps the scene, one JSON line a step:
 *                                        [sha256 of transforms, awake, contacts, frozen], then the moved() checks
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
  const { World } = require(path.join(front, 'world.js'));
  const world = new World();
  const slots = world.addBodies(JSON.parse(fs.readFileSync(scene, 'utf8')));
  const added = Array.from(world.moved(0));
  const n = Number(steps), h = Number(dt);
  let movedOk = added.length === slots.length;
  for (let i = 0; i < n; i++) {
    const before = Float32Array.from(world.transforms);
    world.step(h);
    const t = world.transforms, c = world.counts;
    // moved(stepCount - 1) lists exactly the slots whose pose changed in this step, and after the
    // first step also those added before it (an addition counts as a move in the next step).
    const m = new Set(world.moved(world.stepCount - 1));
    for (let s = 0; s < world.bodyCount; s++) {
      let changed = false;
      for (let k = 0; k < 7; k++) if (t[7 * s + k] !== before[7 * s + k]) changed = true;
      if (changed !== m.has(s) && !(i === 0 && m.has(s))) movedOk = false;
    }
    const hash = crypto.createHash('sha256').update(Buffer.from(t.buffer, t.byteOffset, t.byteLength)).digest('hex');
    process.stdout.write(JSON.stringify([hash, c.awake, c.contacts, c.frozen]) + '\n');
  }
  process.stdout.write(JSON.stringify({ moved: movedOk, stepCount: world.stepCount }) + '\n');
  world.destroy();
} else {
  throw new Error(`run_scene: unknown mode ${mode}`);
}
