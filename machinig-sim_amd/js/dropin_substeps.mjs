// Times the simulator's sub-step loop (src/sim.js:351-449) as a caller of the drop-in
// modules runs it: per sub-step one device call for the engagement (fillEngagement), one for
// the engine's run, and one per stamp (commanded and actual), each answered before the loop
// goes on, as src/sim.js calls them.  The program is one of the recorded simulator runs
// (golden_simx.json: path, tool, engine configuration and wear); the end edges and the wall
// model are left out (no setEndEngagement depth, setWork(null)), so the figure is the cost of
// the calls, not a replay of the run.
//
// usage: node dropin_substeps.mjs golden_simx.json [case] [repetitions]
// prints one JSON line: sub-steps, engine steps and microseconds per sub-step per repetition
import fs from 'fs';
import { Stock, CuttingEngine, fillEngagement, simSchedule, readSubsteps } from './index.js';

const [goldenPath, caseArg = '0', repArg = '3'] = process.argv.slice(2);
const g = JSON.parse(fs.readFileSync(goldenPath, 'utf8'));
const ci = Number(caseArg), reps = Number(repArg);
const cs = g.cases[ci], res = g.results[ci];
const f64 = s => new Float64Array(new Uint8Array(Buffer.from(s, 'base64')).buffer);

const P = res.path, px = f64(P.x), py = f64(P.y), pz = f64(P.z);
const pts = Array.from(px, (x, i) => ({ x, y: py[i], z: pz[i], F: P.F[i] == null ? 0 : P.F[i], S: P.S[i],
  rapid: P.rapid[i] !== 0, line: P.line[i] }));
const sched = simSchedule(pts, { hx: res.hx, N: res.toolN });
const sub = readSubsteps(sched.sub);

const c = res.cfgs[0];
const mode = m => ({ k: m[0], fn: m[1], zeta: m[2] });
const tool = { N: c[1], R: c[2], helixDeg: c[3], type: c[4], fluteLen: c[5], nonlinear: c[6] !== 0, runout: c[7],
  runoutPhase: c[8] };
const modal = { modesX: c[11].map(mode), modesY: c[12].map(mode) };
const k8 = c[13];
const mat = { Ktc: k8[0], Krc: k8[1], Kte: k8[2], Kre: k8[3], Kac: k8[4], Kae: k8[5], mc: k8[6], Ci: k8[7] };
const wear = { Kte: res.wears[0][1], Kre: res.wears[0][2] };
const R = res.toolR, type = res.type === 'ball' ? 'ball' : 'flat';

const times = [];
let steps = 0;
for (let rep = 0; rep <= reps; rep++) {
  const st = cs.stock;
  const stock = new Stock(st.w, st.l, st.h, cs.cell), ideal = new Stock(st.w, st.l, st.h, cs.cell);
  ideal.wallBand = 2.0;
  const e = new CuttingEngine();
  e.configure(tool, modal, mat, c[9], c[10]);
  const sim = { stock, engine: e, activeTool: { R } };
  let fz = 0;
  steps = 0;
  const t0 = process.hrtime.bigint();
  for (const s of sub) {
    const u = [s.ux, s.uy];
    const hmax = s.dxy < 1e-6 ? -Infinity : fillEngagement.call(sim, s.x, s.y, u);
    const ap = Math.min(tool.fluteLen, Math.max(0, hmax - s.z));
    if (ap > 0.01 && !s.rapid) {
      e.setWear(wear);
      e.setEngagement(u, s.fzEff, s.z, ap);
      e.setEndEngagement(0, 0);
      e.setWork(null);
      const want = Math.max(1, Math.round(s.dtm / e.dt));
      const n = Math.max(1, Math.min(want, Math.floor(e.stepsPerTooth / 4)));
      e.run(n);
      steps += n;
      fz = s.fzEff;
    } else {
      e.x *= 0.5; e.y *= 0.5; e.vx *= 0.5; e.vy *= 0.5; e.w *= 0.5; e.wv *= 0.5;
    }
    const dx = (e.x - e.w * e.wnx) * 1e3, dy = (e.y - e.w * e.wny) * 1e3;
    const errVal = Math.hypot(e.x, e.y) * 1e6;
    stock.curLine = ideal.curLine = s.line;
    ideal.stamp(s.x, s.y, s.z, R, type, 0, null, null);
    stock.stamp(s.x + dx, s.y + dy, s.z, R, type, errVal, { ux: u[0], uy: u[1], f: fz * tool.N }, null);
  }
  const us = Number(process.hrtime.bigint() - t0) / 1e3;
  if (rep > 0) times.push(us / sub.length);   // the first run is a warm-up
}
console.log(JSON.stringify({ program: ci, substeps: sub.length, engine_steps: steps, us_per_substep: times }));
