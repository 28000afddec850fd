// Node 12 ES-module loader used only by the fixture generator in this folder.
//
// The reference (bearing-physics, plain ES modules) targets Node >= 20. This
// image has Node 12, which lacks three things the reference's step path
// touches: the `?.` / `??` operators (a handful of sites in save.js and
// joints.js, none on the step path itself), FinalizationRegistry, and
// process.getBuiltinModule (both only in core/jobs.js, the worker pool).
// This loader lowers the two operators in memory, as text, when a module is
// loaded, and serves a one-thread stand-in for core/jobs.js (the reference
// runs on one thread when it has no workers: world.js `_setWorkers`). It
// also lets the generator observe the contact phase: world.js's imports of
// ./contact.js and ./graph.js resolve to the tap modules beside this file,
// which re-export the real modules and record state around their calls.
// Nothing is written to the reference tree.
const REF = 'file:///root/reference/src/';
const HERE = new URL('./', import.meta.url).href;

export async function resolve(specifier, context, defaultResolve) {
  const parent = context.parentURL || '';
  if (parent === REF + 'world.js' && specifier === './contact.js') return { url: HERE + 'tap_contact.mjs' };
  if (parent === REF + 'world.js' && specifier === './graph.js') return { url: HERE + 'tap_graph.mjs' };
  const r = await defaultResolve(specifier, context, defaultResolve);
  if (r.url === REF + 'core/jobs.js') return { url: HERE + 'stub_jobs.mjs' };
  return r;
}

export async function transformSource(source, context, defaultTransformSource) {
  if (!context.url.startsWith(REF)) return defaultTransformSource(source, context, defaultTransformSource);
  let s = typeof source === 'string' ? source : Buffer.from(source).toString('utf8');
  // `x?.f(args)` with a plain name and a call without nested parentheses.
  s = s.replace(/\b([A-Za-z_$][\w$]*)\?\.([A-Za-z_$][\w$]*)\(([^()]*)\)/g, '(($1) == null ? undefined : $1.$2($3))');
  // `x ?? rest;` with a plain name on the left, to the end of the statement.
  s = s.replace(/\b([A-Za-z_$][\w$]*) \?\? ([^;]+);/g, '(($1) != null ? ($1) : ($2));');
  if (/\?\.|\?\?/.test(s.replace(/\/\/.*$/gm, '').replace(/\/\*[\s\S]*?\*\//g, ''))) {
    throw new Error(`lower.mjs: an operator left unlowered in ${context.url}`);
  }
  return { source: s };
}
