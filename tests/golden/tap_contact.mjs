// What world.js imports from the reference's contact.js, re-exported, with
// the phase functions wrapped so the fixture generator can record state
// before and after each call (globalThis.__bmTap, when set). The wrapped
// functions run unchanged.
import * as real from '/root/reference/src/contact.js';

export const { Contacts, PIXEL, WORK, turnedInertias, touched, gapOf, crowded, lockedInverse, makeSoft } = real;

function tap(name, fn) {
  return function (...args) {
    const t = globalThis.__bmTap;
    if (t) t(name, 0, args);
    const r = fn(...args);
    if (t) t(name, 1, args);
    return r;
  };
}

export const narrow = tap('narrow', real.narrow);
export const store = tap('store', real.store);
export const inertias = tap('inertias', real.inertias);
export const prepare = tap('prepare', real.prepare);
export const warmStart = tap('warmStart', real.warmStart);
export const solve = tap('solve', real.solve);
export const restitution = tap('restitution', real.restitution);
export const rest = tap('rest', real.rest);
