// Engagement sampling of the simulator's sub-step on the device-resident Stock:
// Simulator.fillEngagement and Simulator.immersion (src/sim.js:301-348) and the axial depth
// of src/sim.js:388.  fillEngagement and immersion keep the reference's method signatures
// and read this.engine.Hb, this.activeTool.R and this.stock as the reference's do, so a
// Simulator running on this module's Stock takes them as its own methods; engagementBatch
// answers many sub-steps in one launch.
import { native } from './native.js';

export const BINS = 180;
const Q_DOUBLES = 7, RES_BYTES = 40;

// queries: [{cx, cy, ux, uy, R, tipZ?, fluteLen?}] -> [{hmax, ap, imm: {st, ex} | null, Hb}]
export function engagementBatch(stock, queries, wantHb = true) {
  const n = queries.length, q = new Float64Array(Q_DOUBLES * n);
  queries.forEach((e, i) => {
    q.set([e.cx, e.cy, e.ux, e.uy, e.R, e.tipZ === undefined ? 0 : e.tipZ,
      e.fluteLen === undefined ? Infinity : e.fluteLen], Q_DOUBLES * i);
  });
  const Hb = wantHb ? new Float32Array(BINS * n) : null;
  const dv = new DataView(native().stockEngagement(stock.handle, q, Hb));
  return queries.map((_, i) => {
    const o = i * RES_BYTES;
    const imm = dv.getInt32(o + 32, true) !== 0;
    return { hmax: dv.getFloat64(o, true), ap: dv.getFloat64(o + 8, true),
      imm: imm ? { st: dv.getFloat64(o + 16, true), ex: dv.getFloat64(o + 24, true) } : null,
      Hb: Hb ? Hb.subarray(BINS * i, BINS * (i + 1)) : null };
  });
}

// stock height on the cutter periphery per angle bin into this.engine.Hb; returns the highest
export function fillEngagement(cx, cy, u) {
  const r = engagementBatch(this.stock, [{ cx, cy, ux: u[0], uy: u[1], R: this.activeTool.R }])[0];
  this.engine.Hb.set(r.Hb);
  return r.hmax;
}

// engagement arc at the bottom slice: the leading bins standing above tipZ + 0.1, widened by
// half a bin on either side and kept within [0, pi]; null when there are none
export function immersion(tipZ) {
  const Hb = this.engine.Hb, z = tipZ + 0.1, half = Math.PI / BINS;
  let lo = -1, hi = -1;
  for (let b = 0; b < BINS / 2; b++) if (Hb[b] > z) { if (lo < 0) lo = b; hi = b; }
  if (lo < 0) return null;
  const at = b => (b + 0.5) / BINS * 2 * Math.PI;   // bin centre angle
  return { st: Math.max(0, at(lo) - half), ex: Math.min(Math.PI, at(hi) + half) };
}
