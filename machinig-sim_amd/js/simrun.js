// The simulator's sub-step loop (Simulator.step, src/sim.js:351-449) for many programs in
// one device call (mc_sim_run): per sub-step the engagement from the actual stock, the axial
// depth, the engine run for its share of steps, the commanded and the deflected stamp with
// its feed mark; per frame the engine's taken stats and chatter index (Simulator.collect,
// src/sim.js:644-663).  The programs' Stock and CuttingEngine objects are the drop-in ones
// (stock.js, cutting.js): their device state advances, and their host mirrors are refreshed
// as after their own calls.
import { native } from './native.js';
import { MC_FRAME_LAST } from './path.js';

const OUT_BYTES = 80, FRAME_BYTES = 216, TYPE = { flat: 0, ball: 1, bull: 2, profile: 3 };
const STATS = ['n', 'sumF', 'maxF', 'sumFx', 'sumFy', 'sumFz', 'maxFz', 'sumT', 'maxT', 'sumX', 'sumY', 'minX',
  'maxX', 'minY', 'maxY', 'sumZc', 'wZc', 'sumW', 'maxW', 'sumHb', 'sumB', 'sumRb', 'engTeeth'];
const OUT_F = ['dx', 'dy', 'ap', 'errVal', 'vol', 'maxF', 'phiSt', 'phiEx'];

// programs: [{stock (Simulator.stock), ideal (Simulator.ideal), engine (a CuttingEngine
// configured for the program's tool and speed, its setWear applied; the loop starts from the
// state its device calls left), schedule ({sub, n} from simSchedule), tool {R, fluteLen, type: 'flat' | 'ball' |
// 'bull' | 'profile', cornerR, lut, edge (an imported edge {dz, r}, src/assembly.js:31-35)}, rw,
// share, dynBudget, fz, frames (true: frame results), budgetLeft (the frame budget left when the
// call starts inside a frame: after a spindle-speed change re-configured the engine,
// ensureEngine, src/sim.js:252-257)}].
// Returns per program {out: [{dx, dy, ap, errVal, vol, maxF, phiSt, phiEx, nsteps, cut, arc}],
// frames: [{stats, index, amp, once, end}], budgetLeft (the budget left after the last sub-step,
// when budgetLeft was given)}.  The actual stock's removed volume and curLine
// advance as its stamps do; the commanded stock's removed volume is not advanced (the
// simulator reads the actual one).
export function simRun(programs) {
  const args = programs.map(p => {
    const t = p.tool, e = p.engine, mat = e.mat;
    // the engine's setWear reaches the device with its next run: here, this call
    native().engineSetWear(e._h, e.Kte != null ? e.Kte : mat.Kte, e.Kre != null ? e.Kre : mat.Kre);
    if (!(t.type in TYPE)) throw new RangeError('tool type must be flat, ball, bull or profile');
    if (t.type === 'profile' && !t.edge) throw new RangeError('a profile tool needs its edge {dz, r}');
    return { actual: p.stock.handle, ideal: p.ideal.handle, engine: p.engine._h, sub: p.schedule.sub,
      nsub: p.schedule.n, R: t.R, fluteLen: t.fluteLen, rw: p.rw || 0, share: p.share === undefined ? 1 : p.share,
      type: TYPE[t.type], lut: t.type === 'bull' || t.type === 'profile' ? Float32Array.from(t.lut) : null,
      cornerR: t.cornerR || 0, edgeR: t.type === 'profile' ? Float64Array.from(t.edge.r) : null,
      edgeDz: t.type === 'profile' ? t.edge.dz : 0, dynBudget: p.dynBudget || 0, fz: p.fz || 0, frames: !!p.frames,
      budgetLeft: p.budgetLeft === undefined ? null : p.budgetLeft };
  });
  const res = native().simRun(args);
  return programs.map((p, i) => {
    const [ob, fb, left] = res[i], od = new DataView(ob), fd = new DataView(fb);
    const n = ob.byteLength / OUT_BYTES, out = new Array(n);
    let vol = 0;
    for (let k = 0; k < n; k++) {
      const o = k * OUT_BYTES, r = {};
      OUT_F.forEach((f, j) => { r[f] = od.getFloat64(o + 8 * j, true); });
      r.nsteps = od.getInt32(o + 64, true); r.cut = od.getInt32(o + 68, true) !== 0; r.arc = od.getInt32(o + 72, true) !== 0;
      vol += r.vol;
      out[k] = r;
    }
    const frames = [];
    for (let f = 0; f < fb.byteLength / FRAME_BYTES; f++) {
      const o = f * FRAME_BYTES, stats = {};
      STATS.forEach((k, j) => { stats[k] = fd.getFloat64(o + 8 * j, true); });
      frames.push({ stats, index: fd.getFloat64(o + 184, true), amp: fd.getFloat64(o + 192, true),
        once: fd.getFloat64(o + 200, true), end: fd.getInt32(o + 208, true) });
    }
    p.stock.removed += vol;
    if (n > 0) {
      const sd = new DataView(p.schedule.sub);
      p.stock.curLine = sd.getInt32((n - 1) * 96 + 76, true);
      p.ideal.curLine = p.stock.curLine;
    }
    p.stock._changed();
    p.ideal._changed();
    p.engine._sync(native().engineState(p.engine._h), false);
    return left === undefined ? { out, frames } : { out, frames, budgetLeft: left };
  });
}

export { MC_FRAME_LAST };
