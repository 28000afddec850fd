// Writes golden fixtures by running the reference (bearing-physics, mounted
// read-only at /root/reference) on this image's Node 12 through lower.mjs.
// Run from the repository root:
//   python tests/golden/make_fixtures.py
// which writes each scene's JSON (bearing_amd/scenes.py) and calls
//   node --experimental-loader ./tests/golden/lower.mjs tests/golden/make_fixtures.mjs <job.json>
// A job names a scene file, dt, the step count, which steps to record, and the output.
//
// What is recorded (see tests/golden/fixtures.py for the reader):
//   soft   the world's columns and the step's manifolds as contact.js inertias()
//          receives them (after colour() and rest(), src/world.js:956-963), and
//          the same after store() (src/world.js:1031): the phase from the world
//          inverse inertias to the impulses kept on the pairs.
//   trace  for one step, the bodies' velocities (and poses after each move)
//          after every call of that phase, in order.
//   colour the manifold endpoints colour() reads and the colours, order, starts
//          and rows it writes (src/graph.js:105-170).
//   traj   per step, the SHA-256 of world.transforms, of world.previous and of
//          world.moved() since the step before and since seven before, and the
//          transforms themselves every `every` steps.
import { World } from '/root/reference/src/bearing.js';
import { createHash } from 'crypto';
import { readFileSync, writeFileSync } from 'fs';
import { gzipSync } from 'zlib';

const TYPES = new Map([[Float64Array, 'f64'], [Float32Array, 'f32'], [Int32Array, 'i32'], [Uint32Array, 'u32'],
  [Uint8Array, 'u8'], [Int8Array, 'i8']]);

/** A container of named typed arrays: "BMFX0001", u32 header length, JSON header, 8-aligned data. */
function pack(arrays, meta) {
  const entries = [];
  let offset = 0;
  for (const [name, a] of Object.entries(arrays)) {
    const dtype = TYPES.get(a.constructor);
    if (!dtype) throw new Error(`pack: ${name} is not a typed array`);
    entries.push({ name, dtype, length: a.length, offset });
    offset += Math.ceil(a.byteLength / 8) * 8;
  }
  const header = Buffer.from(JSON.stringify({ arrays: entries, meta }));
  const out = Buffer.alloc(12 + header.length + ((8 - ((12 + header.length) % 8)) % 8) + offset);
  out.write('BMFX0001', 0, 'latin1');
  out.writeUInt32LE(header.length, 8);
  header.copy(out, 12);
  const base = 12 + header.length + ((8 - ((12 + header.length) % 8)) % 8);
  let k = 0;
  for (const a of Object.values(arrays)) {
    Buffer.from(a.buffer, a.byteOffset, a.byteLength).copy(out, base + entries[k++].offset);
  }
  return gzipSync(out, { level: 9 });
}

const copy = (a, n) => a.slice(0, n);

const BODY_IN = ['_position', '_rotation', '_velocity', '_angular', '_dims', '_mass', '_invMass', '_inertia',
  '_invInertia', '_invI', '_axes', '_type', '_shape', '_frozen', '_lock', '_jointHead', '_tumbled', '_ceiling',
  '_gravityScale', '_damping', '_angularDamping', '_framed', '_start', '_startQ', '_awake', '_size', '_reach',
  '_friction', '_restitution'];
const BODY_OUT = ['_position', '_rotation', '_velocity', '_angular', '_invI', '_tumbled', '_ceiling', '_start', '_startQ'];
const PER = { _position: 3, _rotation: 4, _velocity: 3, _angular: 3, _dims: 3, _mass: 1, _invMass: 1, _inertia: 3,
  _invInertia: 3, _invI: 6, _axes: 9, _type: 1, _shape: 1, _frozen: 1, _lock: 1, _jointHead: 1, _tumbled: 3,
  _ceiling: 1, _gravityScale: 1, _damping: 1, _angularDamping: 1, _framed: 1, _start: 3, _startQ: 4, _awake: 1,
  _size: 1, _reach: 1, _friction: 1, _restitution: 1 };
// Contacts columns by numbers per manifold (m) or per point slot (4 per manifold).
const C_IN = [['a', 1], ['b', 1], ['pair', 1], ['sub', 1], ['points', 1], ['colour', 1], ['order', 1], ['row', 1],
  ['normal', 3], ['t1', 3], ['t2', 3], ['friction', 1], ['restitution', 1], ['tolerance', 1], ['patch', 1],
  ['rollImpulse1', 1], ['rollImpulse2', 1], ['spinImpulse', 1], ['tangentImpulse1', 1], ['tangentImpulse2', 1],
  ['drift1', 1], ['drift2', 1], ['pid', 4], ['separation', 4], ['point', 12], ['shift', 12], ['normalImpulse', 4],
  ['maxNormalImpulse', 4], ['relativeVelocity', 4], ['gap', 1]];
const C_OUT = [['normalImpulse', 4], ['maxNormalImpulse', 4], ['relativeVelocity', 4], ['gap', 1], ['now', 4],
  ['tangentImpulse1', 1], ['tangentImpulse2', 1], ['rollImpulse1', 1], ['rollImpulse2', 1], ['spinImpulse', 1],
  ['drift1', 1], ['drift2', 1], ['sliding', 1], ['ra', 12], ['rb', 12], ['la', 12], ['lb', 12], ['adjusted', 4],
  ['normalMass', 4], ['jac', 48], ['block', 16], ['lever', 4], ['anchorA', 3], ['anchorB', 3], ['fjac', 24],
  ['tangentMass', 3], ['rollMass1', 1], ['rollMass2', 1], ['spinMass', 1]];
const PAIR_OUT = [['points', 1], ['pid', 4], ['normalImpulse', 4], ['twist', 3], ['friction', 3], ['slip', 3]];

function bodies(world, names, prefix, out) {
  const top = world._slots.top;
  for (const n of names) out[prefix + n.slice(1)] = copy(world[n], top * PER[n]);
}

function contacts(c, cols, prefix, out) {
  for (const [n, k] of cols) out[prefix + n] = copy(c[n], c.count * k);
}

function softPre(world, c, out) {
  bodies(world, BODY_IN, 'in.', out);
  out['in.awakeList'] = copy(world._awakeList, world._awakeCount);
  out['in.gravity'] = Float64Array.from(world._gravity);
  out['in.starts'] = c.starts.slice();
  contacts(c, C_IN, 'in.c.', out);
  return { count: c.count, solved: c.solved, top: world._slots.top, awake: world._awakeCount, dt: world._stepDt,
    substeps: world._substeps, frozenCount: world._frozenCount, step: world._stepCount, pairs: world._pairs.count };
}

function softPost(world, c, out) {
  bodies(world, BODY_OUT, 'out.', out);
  contacts(c, C_OUT, 'out.c.', out);
  const p = world._pairs;
  for (const [n, k] of PAIR_OUT) out['out.pairs.' + n] = copy(p[n], p.count * k);
}

function main() {
  const job = JSON.parse(readFileSync(process.argv[2], 'utf8'));
  const scene = JSON.parse(readFileSync(job.scene, 'utf8'));
  const world = new World({ workers: 0 });
  for (const spec of scene) world.addBody(spec);
  // The bodies as addBody resolved them, before any step: what a binding hands the device world.
  const init = {};
  bodies(world, ['_position', '_rotation', '_velocity', '_angular', '_dims', '_mass', '_invMass', '_inertia',
    '_invInertia', '_type', '_shape', '_framed', '_lock', '_gravityScale', '_damping', '_angularDamping', '_size',
    '_reach', '_friction', '_restitution', '_tumbled', '_ceiling'], 'init.', init);
  // Hull bodies: each one's hull as the step holds it (world._inner, moved to its centre of
  // mass and principal axes), numbered in slot order, and the frame that gives its origin back.
  const top0 = world._slots.top;
  if (Array.from(world._shape.subarray(0, top0)).some((k) => k === 3)) {
    const hullOf = new Int32Array(top0).fill(-1);
    const t = { vertices: [], start: [], loops: [], planes: [], hullV: [0], hullF: [0], hullL: [0] };
    for (let s = 0; s < top0; s++) {
      if (world._shape[s] !== 3) continue;
      const h = world._inner[s];
      hullOf[s] = t.hullV.length - 1;
      for (const x of h.vertices) t.vertices.push(x);
      for (let f = 0; f < h.start.length - 1; f++) t.start.push(h.start[f]);
      for (const x of h.loops) t.loops.push(x);
      for (const x of h.planes) t.planes.push(x);
      t.hullV.push(t.vertices.length / 3);
      t.hullF.push(t.planes.length / 4);
      t.hullL.push(t.loops.length);
    }
    init['init.hull'] = hullOf;
    init['init.offset'] = copy(world._offset, 3 * top0);
    init['init.frame'] = copy(world._frame, 4 * top0);
    init['hulls.vertices'] = Float64Array.from(t.vertices);
    init['hulls.planes'] = Float64Array.from(t.planes);
    for (const k of ['start', 'loops', 'hullV', 'hullF', 'hullL']) init['hulls.' + k] = Int32Array.from(t[k]);
  }
  // Heightfield bodies: each one's field (world._geometry), numbered in slot order.
  if (Array.from(world._shape.subarray(0, top0)).some((k) => k === 6)) {
    const fieldOf = new Int32Array(top0).fill(-1);
    const cols = [], rows = [], spacing = [], heights = [];
    for (let s = 0; s < top0; s++) {
      if (world._shape[s] !== 6) continue;
      const g = world._geometry[s];
      fieldOf[s] = cols.length;
      cols.push(g.columns); rows.push(g.rows); spacing.push(g.spacing);
      for (const h of g.heights) heights.push(h);
    }
    init['init.field'] = fieldOf;
    init['fields.columns'] = Int32Array.from(cols);
    init['fields.rows'] = Int32Array.from(rows);
    init['fields.spacing'] = Float64Array.from(spacing);
    init['fields.heights'] = Float64Array.from(heights);
  }
  const softAt = new Set(job.soft || []), colourAt = new Set(job.colour || []), traceAt = job.trace || -1;
  let rec = null, meta = null, trace = null, traceCount = 0, colourRec = null;
  const snap = (tag) => {
    const top = world._slots.top;
    trace[`t${traceCount}.${tag}.v`] = copy(world._velocity, 3 * top);
    trace[`t${traceCount}.${tag}.w`] = copy(world._angular, 3 * top);
    trace[`t${traceCount}.${tag}.p`] = copy(world._position, 3 * top);
    trace[`t${traceCount}.${tag}.q`] = copy(world._rotation, 4 * top);
    traceCount++;
  };
  const accelerate = World.prototype._accelerate, move = World.prototype._move;
  World.prototype._accelerate = function (h) { accelerate.call(this, h); if (trace) snap('accelerate'); };
  World.prototype._move = function (h) { move.call(this, h); if (trace) snap('move'); };
  globalThis.__bmTap = (name, after, args) => {
    const step = world._stepCount;
    if (name === 'colour' && colourAt.has(step)) {
      const c = args[1];
      if (!after) {
        colourRec = { 'in.a': copy(c.a, c.count), 'in.b': copy(c.b, c.count), 'in.type': copy(world._type, world._slots.top),
          'in.frozen': copy(world._frozen, world._slots.top) };
      } else {
        colourRec['out.colour'] = copy(c.colour, c.count); colourRec['out.order'] = copy(c.order, c.count);
        colourRec['out.row'] = copy(c.row, c.count); colourRec['out.starts'] = c.starts.slice();
        writeFileSync(`${job.out}.colour.${step}.bmfx.gz`, pack(colourRec, { step, count: c.count, solved: c.solved,
          top: world._slots.top, frozenCount: world._frozenCount }));
      }
    }
    if (name === 'inertias' && !after && (softAt.has(step) || step === traceAt)) {
      rec = {};
      meta = softPre(world, args[0]._contacts, rec);
      if (step === traceAt) { trace = {}; traceCount = 0; }
    }
    if (trace && after && (name === 'warmStart' || name === 'solve' || name === 'restitution')) snap(name);
    if (name === 'store' && after && rec !== null) {
      softPost(world, args[1], rec);
      if (softAt.has(step)) writeFileSync(`${job.out}.soft.${step}.bmfx.gz`, pack(rec, meta));
      if (trace) {
        writeFileSync(`${job.out}.trace.${step}.bmfx.gz`, pack(Object.assign({}, rec, trace), Object.assign({ calls: traceCount }, meta)));
        trace = null;
      }
      rec = null;
    }
  };
  const hashes = [], every = job.every || 0, kept = {};
  const counts = { awake: new Int32Array(job.steps), contacts: new Int32Array(job.steps),
    frozen: new Int32Array(job.steps) };
  // world.moved() after every step, as the SHA-256 of its Int32Array (order counts): since the
  // step before (what a renderer drawing every step asks) and since seven steps before; and
  // world.previous. Before the first step, moved(0) (every body added).
  const digest = (a) => createHash('sha256').update(Buffer.from(a.buffer, a.byteOffset, a.byteLength)).digest('hex');
  const moved = { last: [], older: [], previous: [], added: digest(world.moved(0)) };
  for (let s = 1; s <= job.steps; s++) {
    world.step(job.dt);
    const t = world._transforms.subarray(0, 7 * world._slots.top);
    hashes.push(createHash('sha256').update(Buffer.from(t.buffer, t.byteOffset, t.byteLength)).digest('hex'));
    moved.last.push(digest(world.moved(world.stepCount - 1)));
    moved.older.push(digest(world.moved(Math.max(0, world.stepCount - 7))));
    moved.previous.push(digest(world.previous.subarray(0, 7 * world._slots.top)));
    counts.awake[s - 1] = world._awakeCount;
    counts.contacts[s - 1] = world._contacts.count;
    counts.frozen[s - 1] = world._frozenCount;
    if (every > 0 && s % every === 0) kept[`transforms.${s}`] = t.slice();
  }
  const final = {};
  bodies(world, ['_position', '_rotation', '_velocity', '_angular'], 'final.', final);
  writeFileSync(`${job.out}.traj.bmfx.gz`, pack(Object.assign({ 'awake': counts.awake, 'contacts': counts.contacts,
    'frozen': counts.frozen },
    kept, final, init), { steps: job.steps, dt: job.dt, hashes, bodies: scene.length, impacts: world._impacts,
    catches: world._catches, solverSteps: world._solverSteps, moved }));
}

main();
