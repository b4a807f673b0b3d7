// Packing of the reference's plain argument objects into the C records of
// include/mcsim.h (mc_cell 352 B, mc_stamp 48 B, mc_stamp_opt 32 B, mc_lobe_args 256 B),
// little-endian, for the
// addon.
import { modeSet } from './frf.js';
import { native } from './native.js';

export const MC_MAX_MODES = 4, MC_MAX_FLUTES = 8;
export const CELL_BYTES = 352, TRIAL_BYTES = 40, LIMIT_BYTES = 56, STAMP_BYTES = 48, STAMP_OPT_BYTES = 32;
export const LOBE_ARGS_BYTES = 256;

let checked = false;
function checkSizes() {
  if (checked) return;
  const s = native().sizes();
  if (s.mc_cell !== CELL_BYTES || s.mc_trial !== TRIAL_BYTES || s.mc_limit !== LIMIT_BYTES || s.mc_stamp !== STAMP_BYTES ||
      s.mc_stamp_opt !== STAMP_OPT_BYTES || s.mc_lobe_args !== LOBE_ARGS_BYTES)
    throw new Error('mcsim.node record sizes do not match this module: ' + JSON.stringify(s));
  checked = true;
}

const numOr0 = v => (v === null || v === undefined ? 0 : v);

// one chatterAt / timeDomainLimit argument set: tool {N, R, helixDeg, type, fluteLen},
// modal (modeSet conventions), mat {Ktc, Krc, Kac, Kte, Kre, Kae, mc?, Ci?}, S, arc, fz,
// revs, nonlinear
export function writeCell(dv, off, tool, modal, mat, S, arc, fz, revs, nonlinear) {
  const ms = modeSet(modal);
  if (ms.x.length > MC_MAX_MODES || ms.y.length > MC_MAX_MODES)
    throw new RangeError('at most ' + MC_MAX_MODES + ' modes per axis');
  if (!(tool.N >= 1 && tool.N <= MC_MAX_FLUTES)) throw new RangeError('1..' + MC_MAX_FLUTES + ' flutes');
  dv.setInt32(off, tool.N, true);
  dv.setInt32(off + 4, tool.type === 'ball' ? 1 : 0, true);
  dv.setFloat64(off + 8, tool.R, true);
  dv.setFloat64(off + 16, tool.helixDeg, true);
  dv.setFloat64(off + 24, tool.fluteLen, true);
  dv.setFloat64(off + 32, numOr0(tool.runout), true);
  dv.setFloat64(off + 40, numOr0(tool.runoutPhase), true);
  const K = ['Ktc', 'Krc', 'Kac', 'Kte', 'Kre', 'Kae', 'mc', 'Ci'];
  for (let i = 0; i < 8; i++) dv.setFloat64(off + 48 + 8 * i, numOr0(mat[K[i]]), true);
  dv.setInt32(off + 112, ms.x.length, true);
  dv.setInt32(off + 116, ms.y.length, true);
  const put = (base, list) => list.forEach((m, i) => {
    dv.setFloat64(base + 24 * i, m.k, true); dv.setFloat64(base + 24 * i + 8, m.fn, true); dv.setFloat64(base + 24 * i + 16, m.zeta, true);
  });
  put(off + 120, ms.x);
  put(off + 216, ms.y);
  dv.setFloat64(off + 312, S, true);
  dv.setFloat64(off + 320, arc[0], true);
  dv.setFloat64(off + 328, arc[1], true);
  dv.setFloat64(off + 336, fz, true);
  dv.setInt32(off + 344, revs, true);
  dv.setInt32(off + 348, nonlinear ? 1 : 0, true);
}

export function packCells(list) {
  checkSizes();
  const buf = new ArrayBuffer(CELL_BYTES * list.length), dv = new DataView(buf);
  list.forEach((c, i) => writeCell(dv, i * CELL_BYTES, c.tool, c.modal, c.mat, c.S, c.arc, c.fz, c.revs, c.nonlinear));
  return buf;
}

export function readLimits(buf, n) {
  const dv = new DataView(buf), out = [];
  for (let i = 0; i < n; i++) {
    const o = i * LIMIT_BYTES;
    out.push({ alim: dv.getFloat64(o, true), lo: dv.getFloat64(o + 8, true), hi: dv.getFloat64(o + 16, true),
      open: dv.getInt32(o + 24, true) !== 0, trials: dv.getInt32(o + 28, true),
      steps: Number(dv.getBigInt64(o + 32, true)), visits: Number(dv.getBigInt64(o + 40, true)),
      engaged: Number(dv.getBigInt64(o + 48, true)) });
  }
  return out;
}

export function readTrials(buf, n) {
  const dv = new DataView(buf), out = [];
  for (let i = 0; i < n; i++) {
    const o = i * TRIAL_BYTES;
    out.push({ index: dv.getFloat64(o, true), amp: dv.getFloat64(o + 8, true), once: dv.getFloat64(o + 16, true),
      chatter: dv.getInt32(o + 24, true) !== 0, nz: dv.getInt32(o + 28, true), steps: Number(dv.getBigInt64(o + 32, true)) });
  }
  return out;
}

export function packStamps(list) {
  checkSizes();
  const buf = new ArrayBuffer(STAMP_BYTES * list.length), dv = new DataView(buf);
  list.forEach((s, i) => {
    const o = i * STAMP_BYTES;
    dv.setFloat64(o, s.cx, true); dv.setFloat64(o + 8, s.cy, true); dv.setFloat64(o + 16, s.tipZ, true);
    dv.setFloat64(o + 24, s.R, true); dv.setInt32(o + 32, s.type === 'ball' ? 1 : 0, true);
    dv.setInt32(o + 36, s.line | 0, true); dv.setFloat64(o + 40, s.errVal, true);
  });
  return buf;
}

// the mark / lut arguments of the stamps in list: {opts: ArrayBuffer of mc_stamp_opt, luts,
// lutOff, lutLen}, each distinct profile table (by identity) packed once; null when no
// stamp has either
export function packStampOpts(list) {
  if (!list.some(s => s.mark || s.lut)) return null;
  checkSizes();
  const buf = new ArrayBuffer(STAMP_OPT_BYTES * list.length), dv = new DataView(buf);
  const index = new Map(), tabs = [];
  list.forEach((s, i) => {
    const o = i * STAMP_OPT_BYTES, m = s.mark;
    dv.setFloat64(o, m ? m.ux : 0, true); dv.setFloat64(o + 8, m ? m.uy : 0, true);
    dv.setFloat64(o + 16, m && m.f > 0 ? m.f : 0, true);
    let t = -1;
    if (s.lut) {
      if (!index.has(s.lut)) { index.set(s.lut, tabs.length); tabs.push(s.lut); }
      t = index.get(s.lut);
    }
    dv.setInt32(o + 24, t, true); dv.setInt32(o + 28, 0, true);
  });
  const lutOff = new Int32Array(tabs.length), lutLen = new Int32Array(tabs.length);
  let total = 0;
  tabs.forEach((t, i) => { lutOff[i] = total; lutLen[i] = t.length; total += t.length; });
  const luts = new Float32Array(Math.max(1, total));
  tabs.forEach((t, i) => luts.set(t, lutOff[i]));
  return { opts: buf, luts, lutOff, lutLen };
}

// one stabilityLobes parameter set (src/sld.js:23: N, Ktc, Krc, phiSt, phiEx, lobes, rpmMin,
// rpmMax and the modal fields modeSet reads) as an mc_lobe_args record
export function writeLobeArgs(dv, off, p) {
  checkSizes();
  const ms = modeSet(p);
  if (ms.x.length > MC_MAX_MODES || ms.y.length > MC_MAX_MODES)
    throw new RangeError('at most ' + MC_MAX_MODES + ' modes per axis');
  dv.setInt32(off, p.N, true);
  dv.setInt32(off + 4, p.lobes === undefined ? 8 : p.lobes, true);
  const D = [p.Ktc, p.Krc, p.phiSt, p.phiEx, p.rpmMin === undefined ? 1000 : p.rpmMin,
    p.rpmMax === undefined ? 40000 : p.rpmMax];
  for (let i = 0; i < 6; i++) dv.setFloat64(off + 8 + 8 * i, D[i], true);
  dv.setInt32(off + 56, ms.x.length, true);
  dv.setInt32(off + 60, ms.y.length, true);
  const put = (list, at) => list.forEach((m, i) => {
    dv.setFloat64(at + 24 * i, m.k, true);
    dv.setFloat64(at + 24 * i + 8, m.fn, true);
    dv.setFloat64(at + 24 * i + 16, m.zeta, true);
  });
  put(ms.x, off + 64);
  put(ms.y, off + 64 + 24 * MC_MAX_MODES);
}
