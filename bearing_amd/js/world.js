'use strict';
/*
 * World for a JavaScript host, stepped on the MI355X through the N-API addon
 * (bearing_amd/_build/bearing_napi.node, over the C ABI's bm_world_*): the reference's
 * per-step surface, new World(options) (src/world.js:314), addBody (:563), step(dt)
 * (:880), transforms and previous (:547, :554: 7 floats a slot, position then rotation),
 * stepCount (:526) and moved(since) (:771). Bodies are boxes, spheres, capsules, static
 * heightfields and hulls given as built (resolve.js); the world keeps every hull body's
 * moved hull and every heightfield in the tables the device reads. Plain CommonJS that
 * Node 12.22 runs.
 *
 * There is no host fallback: without the addon, or without a GPU, the constructor throws.
 */
const path = require('path');
const { resolve } = require('./resolve');

const ADDON = path.join(__dirname, '..', '_build', 'bearing_napi.node');

function fail(caller, message) {
  throw new Error(`${caller}: ${message}`);
}

/* Hulls (resolve.js's moved hulls, in hull number order) joined into bm_world_set_hulls's layout. */
function joinHulls(hulls) {
  let V = 0, F = 0, L = 0;
  for (const h of hulls) { V += h.vertices.length / 3; F += h.start.length - 1; L += h.loops.length; }
  const n = hulls.length;
  const t = {
    hullV: new Int32Array(n + 1), hullF: new Int32Array(n + 1), hullL: new Int32Array(n + 1),
    vertices: new Float64Array(3 * V), start: new Int32Array(F), loops: new Int32Array(L), planes: new Float64Array(4 * F),
  };
  V = 0; F = 0; L = 0;
  hulls.forEach((h, k) => {
    const faces = h.start.length - 1;
    t.vertices.set(h.vertices, 3 * V);
    t.start.set(Int32Array.from(h.start).subarray(0, faces), F);
    t.loops.set(h.loops, L);
    t.planes.set(h.planes, 4 * F);
    V += h.vertices.length / 3; F += faces; L += h.loops.length;
    t.hullV[k + 1] = V; t.hullF[k + 1] = F; t.hullL[k + 1] = L;
  });
  return t;
}

/* Heightfields ({columns, rows, spacing, heights}, in field number order) joined into bm_world_set_fields's layout. */
function joinFields(fields) {
  const n = fields.length;
  const t = { columns: new Int32Array(n), rows: new Int32Array(n), spacing: new Float64Array(n), offsets: new Float64Array(n + 1) };
  fields.forEach((f, k) => {
    t.columns[k] = f.columns; t.rows[k] = f.rows; t.spacing[k] = f.spacing;
    t.offsets[k + 1] = t.offsets[k] + f.heights.length;
  });
  t.heights = new Float64Array(t.offsets[n]);
  fields.forEach((f, k) => t.heights.set(f.heights, t.offsets[k]));
  return t;
}

class World {
  /* options: {gravity: [x, y, z], substeps, sleeping}, as the reference's constructor. */
  constructor(options) {
    const o = options === undefined ? {} : options;
    if (o === null || typeof o !== 'object') fail('World', 'options must be an object');
    this._native = require(ADDON);
    const gravity = o.gravity === undefined ? null : Float64Array.from(o.gravity);
    const substeps = o.substeps === undefined ? 0 : o.substeps;
    if (!Number.isInteger(substeps) || substeps < 0) fail('World', `substeps must be a whole number, got ${substeps}`);
    this._world = this._native.create(gravity, substeps, o.sleeping === undefined ? true : Boolean(o.sleeping));
    this._stepCount = 0;
    this._top = 0;
    this._counts = null;
    this._transforms = new Float32Array(0);
    this._previous = new Float32Array(0);
    this._movedOut = new Int32Array(0);
    this._hulls = [];
    this._fields = [];
  }

  /* Adds one body; returns its slot. */
  addBody(options) {
    return this.addBodies([options])[0];
  }

  /* Adds bodies in slot order, in one call to the device world; returns their slots. */
  addBodies(list) {
    this._live('addBody');
    if (!Array.isArray(list)) fail('addBody', 'bodies must be an array of options');
    const cols = resolve('addBody', list, this._hulls.length, this._fields.length);
    // The tables first: a body is refused if it names a hull or field the world lacks.
    if (cols.hulls.length > 0) {
      for (const h of cols.hulls) this._hulls.push(h);
      this._native.setHulls(this._world, joinHulls(this._hulls));
    }
    if (cols.fields.length > 0) {
      for (const f of cols.fields) this._fields.push(f);
      this._native.setFields(this._world, joinFields(this._fields));
    }
    this._native.add(this._world, cols);
    const first = this._top;
    this._sync();
    const slots = [];
    for (let s = first; s < this._top; s++) slots.push(s);
    return slots;
  }

  /*
   * One step. The device world counts it only once it has begun changing the world: a step
   * it refuses at the door (dt) leaves stepCount as it was; one that failed part-way is
   * counted, since the world is then part-way through it (and refuses every later step).
   */
  step(dt) {
    this._live('step');
    if (typeof dt !== 'number' || !(dt > 0) || !Number.isFinite(dt)) fail('step', `dt must be a number above 0, got ${dt}`);
    try {
      this._native.step(this._world, dt);
    } catch (e) {
      this._stepCount++;
      this._sync();
      throw e;
    }
    this._stepCount++;
    this._sync();
  }

  get transforms() {
    this._live('transforms');
    return this._transforms;
  }

  get previous() {
    this._live('previous');
    return this._previous;
  }

  get stepCount() {
    this._live('stepCount');
    return this._stepCount;
  }

  get bodyCount() {
    this._live('bodyCount');
    return this._top;
  }

  /* The step's counters: {awake, contacts, pairs, frozen, impacts, catches}. */
  get counts() {
    this._live('counts');
    return this._counts;
  }

  /*
   * The slots of the bodies that moved in any step after step `since`, and always of those
   * that moved in the last one; with bodies added since the last step: the reference's list in
   * the reference's order (the device world keeps its stamps and moved log, src/world.js:771-803).
   * A view the next call or step overwrites.
   */
  moved(since) {
    this._live('moved');
    if (!Number.isInteger(since) || since < 0) fail('moved', `since must be a whole number, 0 or more, got ${since}`);
    if (this._top === 0) return this._movedOut.subarray(0, 0);
    const k = this._native.moved(this._world, since, this._movedOut);
    return this._movedOut.subarray(0, k);
  }

  /*
   * The last step's contact events (src/world.js:690-723): which pairs began touching and which
   * stopped, as fresh plain objects, bodies by slot. Given a slot, only that body's own, each
   * turned so that `a` is that body (the normal then points the other way).
   */
  contacts(slot) {
    this._live('contacts');
    if (slot !== undefined && (!Number.isInteger(slot) || slot < 0 || slot >= this._top)) {
      fail('contacts', `slot must be a body's slot, got ${slot}`);
    }
    const e = this._native.contacts(this._world), out = [];
    for (let i = 0; i < e.length / 10; i++) {
      const r = 10 * i;
      let a = e[r + 1], b = e[r + 2];
      const turned = slot !== undefined && b === slot;
      if (slot !== undefined && a !== slot && !turned) continue;
      if (turned) { b = a; a = slot; }
      if (e[r] === 1) {
        out.push({ type: 'end', a, b });
        continue;
      }
      // 0 - x keeps a zero from becoming -0.
      const k = (x) => (turned ? 0 - x : x);
      out.push({ type: 'begin', a, b, point: [e[r + 3], e[r + 4], e[r + 5]],
        normal: [k(e[r + 6]), k(e[r + 7]), k(e[r + 8])], speed: e[r + 9] });
    }
    return out;
  }

  /* Milliseconds spent so far in each part of the step (the C ABI's bm_world_times). */
  times() {
    this._live('times');
    return this._native.times(this._world);
  }

  destroy() {
    if (this._world !== null) this._native.destroy(this._world);
    this._world = null;
  }

  _live(caller) {
    if (this._world === null) fail(caller, 'the world was destroyed');
  }

  /* Reads the device world's counters and transforms, growing the slot arrays with it. */
  _sync() {
    const top = this._native.read(this._world, null).top;
    if (top !== this._top) {
      const grow = (A, old, per) => { const a = new A(per * top); a.set(old.subarray(0, Math.min(old.length, per * top))); return a; };
      this._transforms = grow(Float32Array, this._transforms, 7);
      this._previous = grow(Float32Array, this._previous, 7);
      this._movedOut = new Int32Array(top);
      this._top = top;
    }
    this._counts = this._native.read(this._world, this._transforms, this._previous);
  }
}

module.exports = { World, resolve, joinHulls, joinFields };
