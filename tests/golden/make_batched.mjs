// Steps a range of independent worlds through the reference (bearing-physics, mounted
// read-only at /root/reference) on this image's Node 12 through lower.mjs, and writes
// what each world's trajectory came to as JSON:
//   node --experimental-loader ./tests/golden/lower.mjs tests/golden/make_batched.mjs <job.json>
// Run by tests/golden/make_fixtures.py (job "batched"), which splits the worlds over
// several of these processes and packs their results into one fixture.
//
// A job: {scenes: a JSON file holding a list of scenes (each a list of addBody options),
// from, to: the range of worlds to step, steps, dt, sample: world numbers whose every step
// is recorded, out: where to write}. Per world: the SHA-256 of world.transforms after the
// last step, the awake, contact and frozen counts then, world._impacts and world._catches;
// per sampled world the same after every step.
import { World } from '/root/reference/src/bearing.js';
import { createHash } from 'crypto';
import { readFileSync, writeFileSync } from 'fs';

function hashOf(world) {
  const t = world._transforms.subarray(0, 7 * world._slots.top);
  return createHash('sha256').update(Buffer.from(t.buffer, t.byteOffset, t.byteLength)).digest('hex');
}

function main() {
  const job = JSON.parse(readFileSync(process.argv[2], 'utf8'));
  const scenes = JSON.parse(readFileSync(job.scenes, 'utf8'));
  const sample = new Set(job.sample || []);
  const out = { from: job.from, to: job.to, final: [], awake: [], contacts: [], frozen: [], impacts: [], catches: [],
    sample: {} };
  for (let i = job.from; i < job.to; i++) {
    const world = new World({ workers: 0 });
    for (const spec of scenes[i]) world.addBody(spec);
    const rec = sample.has(i) ? { hashes: [], awake: [], contacts: [], frozen: [] } : null;
    for (let s = 1; s <= job.steps; s++) {
      world.step(job.dt);
      if (rec !== null) {
        rec.hashes.push(hashOf(world));
        rec.awake.push(world._awakeCount);
        rec.contacts.push(world._contacts.count);
        rec.frozen.push(world._frozenCount);
      }
    }
    out.final.push(hashOf(world));
    out.awake.push(world._awakeCount);
    out.contacts.push(world._contacts.count);
    out.frozen.push(world._frozenCount);
    out.impacts.push(world._impacts);
    out.catches.push(world._catches);
    if (rec !== null) out.sample[i] = rec;
  }
  writeFileSync(job.out, JSON.stringify(out));
}

main();
