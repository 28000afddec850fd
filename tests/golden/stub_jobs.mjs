// One-thread stand-in for the reference's core/jobs.js: no shared memory, one
// core, so World builds no job system and every job runs on the stepping
// thread (world.js `_setWorkers`).
export const sharedMemory = false;
export function cores() { return 1; }
export class Jobs {
  constructor() { throw new Error('stub_jobs: the worker pool is not available under Node 12'); }
}
