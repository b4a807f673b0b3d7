// The path format and the simulator's sub-step schedule through the library (host C++ in
// the reference's arithmetic): discretize (src/gcode.js:435-466) with the reference's
// signature, and simSchedule, the sub-steps Simulator.step visits (src/sim.js:351-381,
// 389-405), packed as the mc_substep records simRun takes.
import { native } from './native.js';

export const MOVE_BYTES = 120, POINT_BYTES = 56, SUBSTEP_BYTES = 96;
export const MC_FRAME_FIRST = 1, MC_FRAME_LAST = 2;
const AX = { x: 0, y: 1, z: 2 };

let checked = false;
function checkSizes() {
  if (checked) return;
  const s = native().sizes();
  if (s.mc_move !== MOVE_BYTES || s.mc_path_point !== POINT_BYTES || s.mc_substep !== SUBSTEP_BYTES)
    throw new Error('mcsim.node path record sizes do not match this module: ' + JSON.stringify(s));
  checked = true;
}

// the parser's moves ({from, to: {x, y, z}, rapid, F, S, line, arc?: {ca, cb, cw, axes, sweep}})
// as mc_move records
export function packMoves(moves) {
  checkSizes();
  const buf = new ArrayBuffer(MOVE_BYTES * moves.length), dv = new DataView(buf);
  moves.forEach((mv, i) => {
    const o = i * MOVE_BYTES;
    ['x', 'y', 'z'].forEach((k, j) => { dv.setFloat64(o + 8 * j, mv.from[k], true); dv.setFloat64(o + 24 + 8 * j, mv.to[k], true); });
    dv.setFloat64(o + 48, mv.F, true); dv.setFloat64(o + 56, mv.S, true);
    dv.setInt32(o + 64, mv.rapid ? 1 : 0, true); dv.setInt32(o + 68, mv.line | 0, true);
    const a = mv.arc;
    dv.setInt32(o + 72, a ? 1 : 0, true);
    if (a) {
      dv.setInt32(o + 76, a.cw ? 1 : 0, true);
      (a.axes || ['x', 'y', 'z']).forEach((k, j) => dv.setInt32(o + 80 + 4 * j, AX[k], true));
      dv.setInt32(o + 92, a.sweep != null ? 1 : 0, true);   // absent (or null in JSON)
      dv.setFloat64(o + 96, a.ca, true); dv.setFloat64(o + 104, a.cb, true);
      dv.setFloat64(o + 112, a.sweep != null ? a.sweep : 0, true);
    }
  });
  return buf;
}

// discretize(moves, maxSeg), src/gcode.js:435: the polyline, each point with the attributes
// of the move whose segment ends at it (the first: the first move's start, as a rapid)
export function discretize(moves, maxSeg = 0.5) {
  if (!moves.length) return [];
  const buf = native().pathDiscretize(packMoves(moves), maxSeg), dv = new DataView(buf);
  const n = buf.byteLength / POINT_BYTES, pts = new Array(n);
  for (let i = 0; i < n; i++) {
    const o = i * POINT_BYTES, mv = moves[dv.getInt32(o + 48, true)];
    pts[i] = { x: dv.getFloat64(o, true), y: dv.getFloat64(o + 8, true), z: dv.getFloat64(o + 16, true),
      rapid: dv.getInt32(o + 40, true) !== 0, F: mv.F, S: mv.S, line: mv.line, tool: mv.tool, wcs: mv.wcs,
      local: mv.local, zero: mv.zero };
  }
  return pts;
}

// the polyline ({x, y, z, rapid, F, S, line}) as mc_path_point records
export function packPoints(pts) {
  checkSizes();
  const buf = new ArrayBuffer(POINT_BYTES * pts.length), dv = new DataView(buf);
  pts.forEach((p, i) => {
    const o = i * POINT_BYTES;
    dv.setFloat64(o, p.x, true); dv.setFloat64(o + 8, p.y, true); dv.setFloat64(o + 16, p.z, true);
    dv.setFloat64(o + 24, p.F, true); dv.setFloat64(o + 32, p.S, true);
    dv.setInt32(o + 40, p.rapid ? 1 : 0, true); dv.setInt32(o + 44, p.line | 0, true); dv.setInt32(o + 48, i, true);
  });
  return buf;
}

// simSchedule(pts, {hx, N, dtReal, speed, feedOverride, maxFrames}): the sub-steps of the
// polyline.  dtReal 0 (the default): every sub-step with an unbounded budget; dtReal > 0:
// successive Simulator.step(dtReal) frames at playback speed `speed`.  Returns {sub
// (ArrayBuffer of mc_substep), n, frames: [[first, count], ...]}.
export function simSchedule(pts, o) {
  const dt = o.dtReal || 0;
  const [sub, fr] = native().simSchedule(packPoints(pts), o.hx, dt, o.speed === undefined ? 1 : o.speed,
    o.feedOverride === undefined ? 1 : o.feedOverride, o.N, o.maxFrames === undefined ? 2147483647 : o.maxFrames);
  const frames = [];
  for (let f = 0; f < fr.length; f += 2) frames.push([fr[f], fr[f + 1]]);
  return { sub, n: sub.byteLength / SUBSTEP_BYTES, frames };
}

const SUB_F = ['x', 'y', 'z', 'ux', 'uy', 'dtm', 'fzEff', 'fza', 'dxy'];

// the mc_substep records of a schedule as objects
export function readSubsteps(buf) {
  const dv = new DataView(buf), n = buf.byteLength / SUBSTEP_BYTES, out = new Array(n);
  for (let i = 0; i < n; i++) {
    const o = i * SUBSTEP_BYTES, r = {};
    SUB_F.forEach((k, j) => { r[k] = dv.getFloat64(o + 8 * j, true); });
    r.rapid = dv.getInt32(o + 72, true) !== 0; r.line = dv.getInt32(o + 76, true);
    r.nSubEst = dv.getFloat64(o + 80, true); r.guard = dv.getInt32(o + 88, true); r.frame = dv.getInt32(o + 92, true);
    out[i] = r;
  }
  return out;
}
