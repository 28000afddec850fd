'use strict';
/*
 * addBody's options resolved into the columns the world keeps by slot (bm_bodies), for
 * boxes and spheres: src/world.js:206-216 (DEFAULTS: a dynamic 1 m box, density 1000,
 * friction 0.5, restitution 0, no damping, gravity scale 1) and src/world.js:2150-2245
 * (_write: mass from density and volume unless a mass is given, the principal inertias of
 * a solid box or ball, the thinnest size and the reach, inverses where positive; a static
 * body has no inverse mass). Plain JavaScript, no addon: it runs without a GPU.
 */

const DEFAULTS = Object.freeze({
  type: 'dynamic', shape: 'box', size: [1, 1, 1], radius: 0.5,
  position: [0, 0, 0], rotation: [0, 0, 0, 1], velocity: [0, 0, 0], angularVelocity: [0, 0, 0],
  density: 1000, friction: 0.5, restitution: 0, damping: 0, angularDamping: 0, gravityScale: 1,
});

/* Options this front-end takes; any other is refused rather than ignored. */
const TAKEN = new Set(Object.keys(DEFAULTS).concat(['mass']));

function fail(caller, message) {
  throw new Error(`${caller}: ${message}`);
}

function vector(caller, name, v, n) {
  if (!Array.isArray(v) && !ArrayBuffer.isView(v)) fail(caller, `${name} must be an array of ${n} numbers`);
  if (v.length !== n) fail(caller, `${name} must have ${n} numbers, got ${v.length}`);
  for (let i = 0; i < n; i++) {
    if (typeof v[i] !== 'number' || !Number.isFinite(v[i])) fail(caller, `${name} must hold finite numbers`);
  }
  return v;
}

/*
 * A rotation as the world keeps it (src/core/check.js:105-116 quat): scaled by its largest
 * magnitude, then to unit length; all zero is refused.
 */
function rotation(caller, q) {
  vector(caller, 'rotation', q, 4);
  const m = Math.max(Math.abs(q[0]), Math.abs(q[1]), Math.abs(q[2]), Math.abs(q[3]));
  if (!(m > 0)) fail(caller, 'rotation must be 4 finite numbers, not all zero');
  const x = q[0] / m, y = q[1] / m, z = q[2] / m, w = q[3] / m;
  const k = 1 / Math.sqrt(x * x + y * y + z * z + w * w);
  return [x * k, y * k, z * k, w * k];
}

function positive(caller, name, x) {
  if (typeof x !== 'number' || !(x > 0) || !Number.isFinite(x)) fail(caller, `${name} must be a number above 0, got ${x}`);
  return x;
}

/*
 * The columns of `list` (addBody options, one entry a body), as the addon's add() takes
 * them: {n, type, shape, lock, position, rotation, velocity, angular, dims, inertia,
 * inv_inertia, mass, inv_mass, gravity_scale, damping, angular_damping, size, reach,
 * friction, restitution}.
 */
function resolve(caller, list) {
  const n = list.length;
  const out = {
    n,
    type: new Uint8Array(n), shape: new Uint8Array(n), lock: new Uint8Array(n),
    position: new Float64Array(3 * n), rotation: new Float64Array(4 * n), velocity: new Float64Array(3 * n),
    angular: new Float64Array(3 * n), dims: new Float64Array(3 * n), inertia: new Float64Array(3 * n),
    inv_inertia: new Float64Array(3 * n), mass: new Float64Array(n), inv_mass: new Float64Array(n),
    gravity_scale: new Float64Array(n), damping: new Float64Array(n), angular_damping: new Float64Array(n),
    size: new Float64Array(n), reach: new Float64Array(n), friction: new Float64Array(n),
    restitution: new Float64Array(n),
  };
  for (let i = 0; i < n; i++) {
    const given = list[i] === undefined ? {} : list[i];
    if (given === null || typeof given !== 'object') fail(caller, 'options must be an object');
    for (const k of Object.keys(given)) {
      if (!TAKEN.has(k)) fail(caller, `${k} is not taken by this world (boxes and spheres; the C ABI takes the rest)`);
    }
    const o = Object.assign({}, DEFAULTS, given);
    if (o.type !== 'dynamic' && o.type !== 'static') fail(caller, `type must be 'dynamic' or 'static', got ${o.type}`);
    const dynamic = o.type === 'dynamic';
    let volume, inertia, size, reach, mass;
    if (o.shape === 'box') {
      const s = vector(caller, 'size', o.size, 3);
      const w = positive(caller, 'size', s[0]), h = positive(caller, 'size', s[1]), d = positive(caller, 'size', s[2]);
      out.dims[3 * i] = w; out.dims[3 * i + 1] = h; out.dims[3 * i + 2] = d;
      volume = w * h * d;
      mass = given.mass !== undefined ? positive(caller, 'mass', given.mass) : positive(caller, 'density', o.density) * volume;
      inertia = [mass * (h * h + d * d) / 12, mass * (w * w + d * d) / 12, mass * (w * w + h * h) / 12];
      size = Math.min(w, h, d);
      reach = 0.5 * Math.sqrt(w * w + h * h + d * d);
    } else if (o.shape === 'sphere') {
      const r = positive(caller, 'radius', o.radius);
      out.dims[3 * i] = r;
      volume = (4 / 3) * Math.PI * r * r * r;
      mass = given.mass !== undefined ? positive(caller, 'mass', given.mass) : positive(caller, 'density', o.density) * volume;
      inertia = [0.4 * mass * r * r, 0.4 * mass * r * r, 0.4 * mass * r * r];
      size = 2 * r;
      reach = r;
    } else {
      fail(caller, `a ${o.shape} is not taken by this front-end (boxes and spheres; the C ABI takes capsules, hulls and heightfields)`);
    }
    out.type[i] = dynamic ? 0 : 1;
    out.shape[i] = o.shape === 'box' ? 0 : 1;
    out.position.set(vector(caller, 'position', o.position, 3), 3 * i);
    out.rotation.set(rotation(caller, o.rotation), 4 * i);
    out.velocity.set(vector(caller, 'velocity', o.velocity, 3), 3 * i);
    out.angular.set(vector(caller, 'angularVelocity', o.angularVelocity, 3), 3 * i);
    for (let a = 0; a < 3; a++) {
      out.inertia[3 * i + a] = inertia[a];
      out.inv_inertia[3 * i + a] = inertia[a] > 0 ? 1 / inertia[a] : 0;
    }
    out.mass[i] = mass;
    out.inv_mass[i] = dynamic ? 1 / mass : 0;
    out.gravity_scale[i] = o.gravityScale;
    out.damping[i] = o.damping;
    out.angular_damping[i] = o.angularDamping;
    out.size[i] = size;
    out.reach[i] = reach;
    out.friction[i] = o.friction;
    out.restitution[i] = o.restitution;
  }
  return out;
}

module.exports = { DEFAULTS, resolve };
