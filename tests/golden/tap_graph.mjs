// world.js's import of the reference's graph.js, with colour() wrapped as in tap_contact.mjs.
import * as real from '/root/reference/src/graph.js';

export const { COLOURS, DIGITS, OVERFLOW, RESTING } = real;

export function colour(...args) {
  const t = globalThis.__bmTap;
  if (t) t('colour', 0, args);
  const r = real.colour(...args);
  if (t) t('colour', 1, args);
  return r;
}
