// Runs the REFERENCE implementation (read from its source tree at run time, down-levelled
// in memory by downlevel_loader.mjs) on the cases in a JSON file and writes its answers.
// Used only in the build container by make_golden.py; the fixtures it writes are data.
//
// usage: node --experimental-loader ./downlevel_loader.mjs run_reference.mjs cases.json out.json
import fs from 'fs';
import zlib from 'zlib';

const REF = process.env.MC_REFERENCE_SRC || '/root/reference/src/';
const [casesPath, outPath] = process.argv.slice(2);
const cases = JSON.parse(fs.readFileSync(casesPath, 'utf8'));

let CE, NB;
async function main() {
const { CuttingEngine, BINS } = await import(REF + 'cutting.js');
const { chatterAt, timeDomainLimit } = await import(REF + 'stabtd.js');
const { stabilityLobes, alimAt, directionalCoeffs } = await import(REF + 'sld.js');
const { Stock } = await import(REF + 'stock.js');
const { toolModel } = await import(REF + 'tool.js');
const { tipLut } = await import(REF + 'assembly.js');
CE = CuttingEngine; NB = BINS;

const b64f32 = a => Buffer.from(new Float32Array(a).buffer).toString('base64');
const b64i32 = a => Buffer.from(new Int32Array(a).buffer).toString('base64');
const b64f64 = a => Buffer.from(new Float64Array(a).buffer).toString('base64');
// deflated (zlib) little-endian bytes, for the larger maps
const z32 = (T, a) => zlib.deflateSync(Buffer.from(new T(a).buffer)).toString('base64');

// the engine chatterAt builds (src/stabtd.js:15-20), kept open to read its chatter index
function chatterEngine(c, ap) {
  const e = new CE();
  const BINS = NB;
  e.configure({ N: c.tool.N, R: c.tool.R, helixDeg: c.tool.helixDeg, type: c.tool.type,
    fluteLen: Math.max(c.tool.fluteLen, ap + 1), runout: 0, nonlinear: c.nonlinear }, c.modal, c.mat, c.S);
  for (let b = 0; b < BINS; b++) { const phi = (b + 0.5) / BINS * 2 * Math.PI; e.Hb[b] = phi >= c.arc[0] && phi <= c.arc[1] ? 0 : -1e3; }
  e.setEngagement([1, 0], c.fz, -ap, ap);
  e.vx = 0.01;
  return e;
}

const out = {};
const t0 = Date.now();

if (cases.chatter) {
  out.chatter = cases.chatter.map(({ cell: c, ap }) => {
    const e = chatterEngine(c, ap);
    e.run(e.stepsPerTooth * c.tool.N * c.revs);
    const ci = e.chatterIndex();
    const verdict = chatterAt(c.tool, c.modal, c.mat, c.S, c.arc, c.fz, ap, { revs: c.revs, nonlinear: c.nonlinear });
    return { index: ci.index, amp: ci.amp, once: ci.once === undefined ? 0 : ci.once, chatter: verdict ? 1 : 0,
      nz: e.nz, stepsPerTooth: e.stepsPerTooth, decim: e.decim };
  });
  console.error('chatter', out.chatter.length, (Date.now() - t0) / 1e3, 's');
}

if (cases.tdl) {
  out.tdl = cases.tdl.map(({ cell: c, opts }) => {
    const r = timeDomainLimit(c.tool, c.modal, c.mat, c.S, c.arc, c.fz, { ...opts, revs: c.revs, nonlinear: c.nonlinear });
    return { alim: r.alim, lo: r.bracket[0], hi: r.bracket[1] === Infinity ? 'inf' : r.bracket[1] };
  });
  console.error('tdl', out.tdl.length, (Date.now() - t0) / 1e3, 's');
}

if (cases.trace) {
  out.trace = cases.trace.map(({ cell: c, ap, steps }) => {
    const e = chatterEngine(c, ap);
    const Fx = [], Fy = [], x = [], y = [];
    e.takeStats();
    for (let s = 0; s < steps; s++) {
      e.run(1);
      const st = e.takeStats();
      Fx.push(st.sumFx); Fy.push(st.sumFy); x.push(e.x); y.push(e.y);
    }
    return { Fx: b64f64(Fx), Fy: b64f64(Fy), x: b64f64(x), y: b64f64(y) };
  });
  console.error('trace', (Date.now() - t0) / 1e3, 's');
}

const num = v => v === -Infinity ? '-inf' : v === Infinity ? 'inf' : v;
const rect = d => [d.i0, d.j0, d.i1, d.j1].map(num);
if (cases.stock) {
  out.stock = cases.stock.map(({ w, l, h, cell, wallBand, stamps, samples, tools, queries, clearAt, walls }) => {
    const c = { walls };
    const s = new Stock(w, l, h, cell);
    if (wallBand) s.wallBand = wallBand;
    const luts = (tools || []).map(tipLut);
    const vol = [], dirty = [], disk = [];
    const ask = () => { if (queries) for (const [x, y, r, a] of queries) {
      const m = s.maxInDisk(x, y, r, a === '-inf' ? -Infinity : a); disk.push([num(m.h), m.x, m.y]); } };
    stamps.forEach((st, q) => {
      if (clearAt !== undefined && q === clearAt) { dirty.push(rect(s.dirty)); ask(); s.clearDirty(); }
      s.curLine = st.line;
      vol.push(s.stamp(st.cx, st.cy, st.tipZ, st.R, st.type, st.errVal, st.mark || null,
        st.lut === undefined ? null : luts[st.lut]));
    });
    dirty.push(rect(s.dirty));
    ask();
    // wall-error marks (markWall) after a clearDirty: the map and the rectangle they leave
    let wallDirty = null;
    if (c.walls) {
      s.clearDirty();
      for (const [x, y, R, nx, ny, v] of c.walls) s.markWall(x, y, R, nx, ny, v);
      wallDirty = rect(s.dirty);
    }
    const bmax = [];
    for (let bj = 0; bj < s.bny; bj++) for (let bi = 0; bi < s.bnx; bi++) bmax.push(s.blockMax(bi, bj));
    const smp = samples.map(([x, y]) => num(s.sample(x, y)));
    return { nx: s.nx, ny: s.ny, hx: s.hx, hy: s.hy, H: b64f32(s.H), err: b64f32(s.err), wallD: b64f32(s.wallD),
      lineCut: b64i32(s.lineCut), lineWall: b64i32(s.lineWall), markS: b64f32(s.markS), markF: b64f32(s.markF),
      vol, removed: s.removed, sample: smp, dirty, disk, bmax: b64f32(bmax), luts: luts.map(b64f32),
      wallErr: b64f32(s.wallErr), wallDirty };
  });
  console.error('stock', (Date.now() - t0) / 1e3, 's');
}

if (cases.lobes) {
  out.lobes = cases.lobes.map(({ params, speeds }) => {
    const lob = stabilityLobes(params);
    const alim = speeds.map(S => { const a = alimAt(lob, S); return a === Infinity ? 'inf' : a; });
    const dc = directionalCoeffs(params.Krc / params.Ktc, params.phiSt, params.phiEx);
    return { lobes: lob.map(L => ({ rpm: b64f64(L.map(p => p.rpm)), alim: b64f64(L.map(p => p.alim)) })), alimAt: alim,
      dc: [dc.xx, dc.xy, dc.yx, dc.yy] };
  });
}

// the simulator's engagement sampling, called as Simulator methods on a stand-in holding
// just the fields they read (engine.Hb, activeTool.R, stock)
// a long stamp batch read from a binary file (mc_stamp records: cx, cy, tipZ, R, type, line,
// errVal; 48 bytes, little-endian), applied in order; the answer is each final map's SHA-256
// and the removed volume (the maps of a 512 x 512 stock are too large to keep as data)
if (cases.zigzag) {
  const crypto = await import('crypto');
  const sha = a => crypto.createHash('sha256').update(Buffer.from(a.buffer, a.byteOffset, a.byteLength)).digest('hex');
  out.zigzag = cases.zigzag.map(({ w, l, h, cell, stampsFile, n }) => {
    const s = new Stock(w, l, h, cell), b = fs.readFileSync(stampsFile);
    const dv = new DataView(b.buffer, b.byteOffset, b.byteLength);
    const t1 = Date.now();
    for (let q = 0; q < n; q++) {
      const o = 48 * q;
      s.curLine = dv.getInt32(o + 36, true);
      s.stamp(dv.getFloat64(o, true), dv.getFloat64(o + 8, true), dv.getFloat64(o + 16, true), dv.getFloat64(o + 24, true),
        dv.getInt32(o + 32, true) === 1 ? 'ball' : 'flat', dv.getFloat64(o + 40, true), null, null);
    }
    const ms = Date.now() - t1;
    return { nx: s.nx, ny: s.ny, n, H: sha(s.H), err: sha(s.err), wallD: sha(s.wallD), lineCut: sha(s.lineCut),
      lineWall: sha(s.lineWall), markS: sha(s.markS), markF: sha(s.markF), removed: s.removed, node_ms: ms };
  });
  console.error('zigzag', (Date.now() - t0) / 1e3, 's');
}

if (cases.engage) {
  const { Simulator } = await import(REF + 'sim.js');
  out.engage = cases.engage.map(({ stock: si, queries }) => {
    const c = cases.stock[si];
    const s = new Stock(c.w, c.l, c.h, c.cell);
    if (c.wallBand) s.wallBand = c.wallBand;
    const luts = (c.tools || []).map(tipLut);
    for (const st of c.stamps) {
      s.curLine = st.line;
      s.stamp(st.cx, st.cy, st.tipZ, st.R, st.type, st.errVal, st.mark || null, st.lut === undefined ? null : luts[st.lut]);
    }
    const Hb = [], hmax = [], imm = [];
    for (const q of queries) {
      const me = { engine: { Hb: new Float32Array(NB) }, activeTool: { R: q.R }, stock: s };
      hmax.push(num(Simulator.prototype.fillEngagement.call(me, q.cx, q.cy, [q.ux, q.uy])));
      Hb.push(...me.engine.Hb);
      const m = Simulator.prototype.immersion.call(me, q.tipZ);
      imm.push(m === null ? null : [m.st, m.ex]);
    }
    return { Hb: b64f32(Hb), hmax, imm };
  });
}

// whole simulator runs (one unbounded frame at a time), recording every stamp it issues to
// the actual and the commanded stock and every engagement sample it takes
if (cases.sim) {
  const { Simulator } = await import(REF + 'sim.js');
  out.sim = cases.sim.map(({ program, tool, stock, cell }) => {
    const sim = new Simulator();
    sim.cfg.cell = cell;
    Object.assign(sim.cfg.tool, tool);
    sim.load(program, { stock, origin: { x: 0, y: 0, z: 0 } });
    const rec = { actual: [], ideal: [], engage: [] };
    const tabs = [];
    const lutId = l => { if (!l) return -1; let i = tabs.indexOf(l); if (i < 0) { tabs.push(l); i = tabs.length - 1; } return i; };
    const wrap = (st, out) => {
      const f = st.stamp.bind(st);
      st.stamp = (cx, cy, tipZ, R, type, errVal, mark, lut) => {
        const v = f(cx, cy, tipZ, R, type, errVal, mark, lut);
        out.push([cx, cy, tipZ, R, type === 'ball' ? 1 : 0, errVal, mark ? mark.ux : 0, mark ? mark.uy : 0,
          mark ? mark.f : 0, lutId(lut), st.curLine, v]);
        return v;
      };
    };
    wrap(sim.stock, rec.actual); wrap(sim.ideal, rec.ideal);
    const fill = sim.fillEngagement.bind(sim);
    sim.fillEngagement = (cx, cy, u) => {
      const h = fill(cx, cy, u);
      rec.engage.push([rec.actual.length, cx, cy, u[0], u[1], sim.activeTool.R, num(h)]);
      return h;
    };
    let frames = 0;
    while (!sim.done && frames < 100) { sim.step(1e6); frames++; }
    const maps = s => ({ H: z32(Float32Array, s.H), err: z32(Float32Array, s.err), wallD: z32(Float32Array, s.wallD),
      lineCut: z32(Int32Array, s.lineCut), lineWall: z32(Int32Array, s.lineWall), markS: z32(Float32Array, s.markS),
      markF: z32(Float32Array, s.markF), removed: s.removed });
    const col = (rows, k) => rows.map(r => r[k]);
    const pack = rows => ({ cx: b64f64(col(rows, 0)), cy: b64f64(col(rows, 1)), tipZ: b64f64(col(rows, 2)),
      R: b64f64(col(rows, 3)), type: col(rows, 4), errVal: b64f64(col(rows, 5)), ux: b64f64(col(rows, 6)),
      uy: b64f64(col(rows, 7)), f: b64f64(col(rows, 8)), lut: col(rows, 9), line: col(rows, 10), vol: b64f64(col(rows, 11)) });
    const P = sim.path;
    return { frames, nx: sim.stock.nx, ny: sim.stock.ny, hx: sim.stock.hx, hy: sim.stock.hy, zBottom: sim.stock.zBottom,
      path: { x: b64f64(P.map(p => p.x)), y: b64f64(P.map(p => p.y)), z: b64f64(P.map(p => p.z)),
        rapid: P.map(p => (p.rapid ? 1 : 0)), line: P.map(p => p.line) },
      actual: pack(rec.actual), ideal: pack(rec.ideal), luts: tabs.map(b64f32),
      engage: rec.engage, stock: maps(sim.stock), commanded: maps(sim.ideal) };
  });
}

// whole simulator runs with a rigid workpiece and the linear force law, for the device
// sub-step pipeline: the path with its feeds and speeds, the engine's configure and wear
// arguments, every stamp issued to the actual and the commanded stock, and the final maps
if (cases.simx) {
  const { Simulator } = await import(REF + 'sim.js');
  const { modeSet } = await import(REF + 'frf.js');
  out.simx = cases.simx.map(({ program, tool, stock, cell }, ci) => {
    const tc = Date.now();
    const sim = new Simulator();
    sim.cfg.cell = cell;
    sim.cfg.workpieceFlex = false;
    sim.cfg.cuttingNonlinear = false;
    // an imported edge profile keeps its radii as float32, as an imported mesh gives them
    if (tool.edge) tool = { ...tool, edge: { dz: tool.edge.dz, r: Float32Array.from(tool.edge.r) } };
    Object.assign(sim.cfg.tool, tool);
    sim.load(program, { stock, origin: { x: 0, y: 0, z: 0 } });
    const rec = { actual: [], ideal: [] };
    const wrap = (st, o) => {
      const f = st.stamp.bind(st);
      st.stamp = (cx, cy, tipZ, R, type, errVal, mark, lut) => {
        const v = f(cx, cy, tipZ, R, type, errVal, mark, lut);
        o.push([cx, cy, tipZ, R, type === 'ball' ? 1 : 0, errVal, mark ? mark.f : 0, st.curLine, v]);
        return v;
      };
    };
    wrap(sim.stock, rec.actual); wrap(sim.ideal, rec.ideal);
    // the engagement arc the simulator shows (src/sim.js:421-423) at each actual stamp
    const darc = [];
    const fa = sim.stock.stamp;
    sim.stock.stamp = (...a) => { darc.push(sim.display.phiSt, sim.display.phiEx); return fa(...a); };
    const e = sim.engine, cfgs = [], wears = [], runs = [];
    const fc = e.configure.bind(e), fw = e.setWear.bind(e), fr = e.run.bind(e);
    e.configure = (t, modal, mat, S, fnMax = 0) => {
      fc(t, modal, mat, S, fnMax);
      const ms = modeSet(modal), md = L => L.map(m => [m.k, m.fn, m.zeta]);
      cfgs.push([rec.actual.length, t.N, t.R, t.helixDeg, t.type, t.fluteLen, t.nonlinear === false ? 0 : 1, t.runout || 0,
        t.runoutPhase || 0, S, fnMax, md(ms.x), md(ms.y),
        [mat.Ktc, mat.Krc, mat.Kte, mat.Kre, mat.Kac, mat.Kae, mat.mc != null ? mat.mc : 0, mat.Ci != null ? mat.Ci : 0]]);
    };
    e.setWear = edge => {
      fw(edge);
      const last = wears[wears.length - 1];
      if (!last || last[1] !== edge.Kte || last[2] !== edge.Kre) wears.push([rec.actual.length, edge.Kte, edge.Kre]);
    };
    e.run = n => { fr(n); runs.push([rec.actual.length, n]); };
    let frames = 0;
    while (!sim.done && frames < 100) { sim.step(1e6); frames++; }
    console.error('simx case', ci, (Date.now() - tc) / 1e3, 's');
    const maps = s => ({ H: z32(Float32Array, s.H), err: z32(Float32Array, s.err), wallD: z32(Float32Array, s.wallD),
      lineCut: z32(Int32Array, s.lineCut), lineWall: z32(Int32Array, s.lineWall), markS: z32(Float32Array, s.markS),
      markF: z32(Float32Array, s.markF), removed: s.removed });
    const col = (rows, k) => rows.map(r => r[k]);
    const pack = rows => ({ cx: b64f64(col(rows, 0)), cy: b64f64(col(rows, 1)), tipZ: b64f64(col(rows, 2)),
      R: b64f64(col(rows, 3)), type: col(rows, 4), errVal: b64f64(col(rows, 5)), f: b64f64(col(rows, 6)),
      line: col(rows, 7), vol: b64f64(col(rows, 8)) });
    const P = sim.path, t = sim.activeTool;
    const clr = t.clearanceDeg != null ? t.clearanceDeg : (sim.cfg.tool.clearanceDeg != null ? sim.cfg.tool.clearanceDeg : 10);
    return { frames, darc: b64f64(darc), nx: sim.stock.nx, ny: sim.stock.ny, hx: sim.stock.hx, hy: sim.stock.hy, zBottom: sim.stock.zBottom,
      toolR: t.R, toolN: t.N, fluteLen: t.fluteLen, type: t.type, VB: sim.VB || 0, clearanceDeg: clr,
      path: { x: b64f64(P.map(p => p.x)), y: b64f64(P.map(p => p.y)), z: b64f64(P.map(p => p.z)),
        rapid: P.map(p => (p.rapid ? 1 : 0)), line: P.map(p => p.line), F: P.map(p => p.F), S: P.map(p => p.S) },
      cfgs, wears, runs, actual: pack(rec.actual), ideal: pack(rec.ideal), stock: maps(sim.stock), commanded: maps(sim.ideal) };
  });
  console.error('simx', (Date.now() - t0) / 1e3, 's');
}

// Simulator runs with the thin-wall model on (workpieceFlex, the simulator's default): what
// simx records, plus every setWork argument the step loop passes (wallAt's answer, src/sim.js:407),
// every markWall call on the actual stock (:443) and the final wall-error map; and, at chosen
// wallAt calls, the call's whole input (the stock's heights, the tool position, Hb, the feed
// direction, the load height, the material and the tool's static stiffness) with the reference
// wallAt's own answer on it, so that wallAt is pinned call by call as well
if (cases.simw) {
  const { Simulator } = await import(REF + 'sim.js');
  const { modeSet } = await import(REF + 'frf.js');
  const { wallAt } = await import(REF + 'wall.js');
  const WF = ['k', 'm', 'fn', 'zeta', 'nx', 'ny', 't', 'tRoot', 'hWall', 'a', 'bEff', 'base', 'cx', 'cy', 'I'];
  out.simw = cases.simw.map(({ program, tool, stock, cell, nonlinear, every, maxSnaps }, ci) => {
    const tc = Date.now();
    const sim = new Simulator();
    sim.cfg.cell = cell;
    sim.cfg.workpieceFlex = true;
    sim.cfg.cuttingNonlinear = nonlinear;
    Object.assign(sim.cfg.tool, tool);
    sim.load(program, { stock, origin: { x: 0, y: 0, z: 0 } });
    const rec = { actual: [], ideal: [] };
    const wrap = (st, o) => {
      const f = st.stamp.bind(st);
      st.stamp = (cx, cy, tipZ, R, type, errVal, mark, lut) => {
        const v = f(cx, cy, tipZ, R, type, errVal, mark, lut);
        o.push([cx, cy, tipZ, R, type === 'ball' ? 1 : 0, errVal, mark ? mark.f : 0, st.curLine, v]);
        return v;
      };
    };
    wrap(sim.stock, rec.actual); wrap(sim.ideal, rec.ideal);
    const marks = [];
    const fm = sim.stock.markWall.bind(sim.stock);
    sim.stock.markWall = (cx, cy, R, nx, ny, val) => { marks.push([rec.actual.length, cx, cy, R, nx, ny, val]); fm(cx, cy, R, nx, ny, val); };
    let px = 0, py = 0, pz = 0, pap = 0, found = 0, none = 0;
    const fe = sim.fillEngagement.bind(sim);
    sim.fillEngagement = (x, y, u) => { px = x; py = y; return fe(x, y, u); };
    const e = sim.engine, cfgs = [], wears = [], runs = [], works = [], snaps = [];
    const fc = e.configure.bind(e), fw = e.setWear.bind(e), fr = e.run.bind(e), fs = e.setEngagement.bind(e),
      fk = e.setWork.bind(e);
    let inCfg = false;
    e.configure = (t, modal, mat, S, fnMax = 0) => {
      inCfg = true;
      fc(t, modal, mat, S, fnMax);
      inCfg = false;
      const ms = modeSet(modal), md = L => L.map(m => [m.k, m.fn, m.zeta]);
      cfgs.push([rec.actual.length, t.N, t.R, t.helixDeg, t.type, t.fluteLen, t.nonlinear === false ? 0 : 1, t.runout || 0,
        t.runoutPhase || 0, S, fnMax, md(ms.x), md(ms.y),
        [mat.Ktc, mat.Krc, mat.Kte, mat.Kre, mat.Kac, mat.Kae, mat.mc != null ? mat.mc : 0, mat.Ci != null ? mat.Ci : 0]]);
    };
    e.setWear = edge => {
      fw(edge);
      const last = wears[wears.length - 1];
      if (!last || last[1] !== edge.Kte || last[2] !== edge.Kre) wears.push([rec.actual.length, edge.Kte, edge.Kre]);
    };
    e.setEngagement = (u, fz, tipZ, ap) => { pz = tipZ; pap = ap; fs(u, fz, tipZ, ap); };
    e.setWork = w => {
      if (!inCfg) {
        works.push(w ? [rec.actual.length, ...WF.map(k => w[k])] : [rec.actual.length]);
        // the same call again on the same state, recorded with its input: every `every`-th call
        // that finds a wall, and a few that find none
        const pick = pap > 0.01 && (w ? found++ % every === 0 && snaps.length < maxSnaps : none++ % 97 === 5 && none < 400);
        if (pick) {
          const st = sim.stock, kTool = sim.toolStaticK(), zLoad = pz + pap / 2;
          const again = wallAt(st, { x: px, y: py, z: pz }, sim.activeTool.R, e.Hb, sim.u, zLoad, sim.mat, kTool);
          if (JSON.stringify(again) !== JSON.stringify(w)) throw new Error('wallAt replay differs');
          snaps.push({ sub: rec.actual.length, x: px, y: py, z: pz, R: sim.activeTool.R, ux: sim.u[0], uy: sim.u[1], zLoad,
            E: sim.mat.E, rho: sim.mat.rho, kTool, Hb: b64f32(e.Hb), H: z32(Float32Array, st.H),
            wall: again ? WF.map(k => again[k]) : null });
        }
      }
      fk(w);
    };
    e.run = n => { fr(n); runs.push([rec.actual.length, n]); };
    let frames = 0;
    while (!sim.done && frames < 100) { sim.step(1e6); frames++; }
    console.error('simw case', ci, (Date.now() - tc) / 1e3, 's', works.filter(w => w.length > 1).length, 'walls');
    const maps = s => ({ H: z32(Float32Array, s.H), err: z32(Float32Array, s.err), wallD: z32(Float32Array, s.wallD),
      lineCut: z32(Int32Array, s.lineCut), lineWall: z32(Int32Array, s.lineWall), markS: z32(Float32Array, s.markS),
      markF: z32(Float32Array, s.markF), wallErr: z32(Float32Array, s.wallErr), removed: s.removed });
    const col = (rows, k) => rows.map(r => r[k]);
    const pack = rows => ({ cx: b64f64(col(rows, 0)), cy: b64f64(col(rows, 1)), tipZ: b64f64(col(rows, 2)),
      R: b64f64(col(rows, 3)), type: col(rows, 4), errVal: b64f64(col(rows, 5)), f: b64f64(col(rows, 6)),
      line: col(rows, 7), vol: b64f64(col(rows, 8)) });
    const P = sim.path, t = sim.activeTool;
    const clr = t.clearanceDeg != null ? t.clearanceDeg : (sim.cfg.tool.clearanceDeg != null ? sim.cfg.tool.clearanceDeg : 10);
    return { frames, nx: sim.stock.nx, ny: sim.stock.ny, hx: sim.stock.hx, hy: sim.stock.hy, zBottom: sim.stock.zBottom,
      toolR: t.R, toolN: t.N, fluteLen: t.fluteLen, type: t.type, VB: sim.VB || 0, clearanceDeg: clr,
      E: sim.mat.E, rho: sim.mat.rho, kTool: sim.toolStaticK(), wallFields: WF,
      path: { x: b64f64(P.map(p => p.x)), y: b64f64(P.map(p => p.y)), z: b64f64(P.map(p => p.z)),
        rapid: P.map(p => (p.rapid ? 1 : 0)), line: P.map(p => p.line), F: P.map(p => p.F), S: P.map(p => p.S) },
      cfgs, wears, runs, snaps, actual: pack(rec.actual),
      // setWork calls: the sub-step, 1 when the argument is a wall, and its fields (wallFields)
      works: { sub: works.map(w => w[0]), wall: works.map(w => (w.length > 1 ? 1 : 0)),
        fields: b64f64([].concat(...works.filter(w => w.length > 1).map(w => w.slice(1)))) },
      // markWall calls: the sub-step and (cx, cy, R, nx, ny, val)
      marks: { sub: marks.map(m => m[0]), args: b64f64([].concat(...marks.map(m => m.slice(1)))) }, ideal: pack(rec.ideal), stock: maps(sim.stock),
      commanded: maps(sim.ideal) };
  });
  console.error('simw', (Date.now() - t0) / 1e3, 's');
}

// whole simulator runs with real-time frames (a rigid workpiece, the linear force law): the
// sub-steps each frame of sim.step(dtReal) takes at playback speed `speed` under the dynamics
// budget, every stamp, every engine run, and per frame (collect) the engine's taken stats,
// chatter index, the display's chatter state, the flank wear before and after, and at the
// end the machining report (LineReport.pack)
if (cases.report) {
  const { Simulator } = await import(REF + 'sim.js');
  const { modeSet } = await import(REF + 'frf.js');
  const ST = ['n', 'sumF', 'maxF', 'sumFx', 'sumFy', 'sumFz', 'maxFz', 'sumT', 'maxT', 'sumX', 'sumY', 'minX',
    'maxX', 'minY', 'maxY', 'sumZc', 'wZc', 'sumW', 'maxW', 'sumHb', 'sumB', 'sumRb', 'engTeeth'];
  out.report = cases.report.map(({ program, tool, stock, cell, speed, dtReal, dynBudget }) => {
    const sim = new Simulator();
    sim.cfg.cell = cell;
    sim.cfg.workpieceFlex = false;
    sim.cfg.cuttingNonlinear = false;
    if (dynBudget) sim.cfg.dynBudget = dynBudget;
    Object.assign(sim.cfg.tool, tool);
    sim.load(program, { stock, origin: { x: 0, y: 0, z: 0 } });
    sim.speed = speed;
    const rec = { actual: [], ideal: [] };
    const wrap = (st, o) => {
      const f = st.stamp.bind(st);
      st.stamp = (cx, cy, tipZ, R, type, errVal, mark, lut) => {
        const v = f(cx, cy, tipZ, R, type, errVal, mark, lut);
        o.push([cx, cy, tipZ, R, type === 'ball' ? 1 : 0, errVal, mark ? mark.f : 0, st.curLine, v]);
        return v;
      };
    };
    wrap(sim.stock, rec.actual); wrap(sim.ideal, rec.ideal);
    const darc = [];
    const fa = sim.stock.stamp;
    sim.stock.stamp = (...a) => { darc.push(sim.display.phiSt, sim.display.phiEx); return fa(...a); };
    const e = sim.engine, cfgs = [], wears = [], runs = [];
    const fc = e.configure.bind(e), fw = e.setWear.bind(e), fr = e.run.bind(e);
    e.configure = (t, modal, mat, S, fnMax = 0) => {
      fc(t, modal, mat, S, fnMax);
      const ms = modeSet(modal), md = L => L.map(m => [m.k, m.fn, m.zeta]);
      cfgs.push([rec.actual.length, t.N, t.R, t.helixDeg, t.type, t.fluteLen, t.nonlinear === false ? 0 : 1, t.runout || 0,
        t.runoutPhase || 0, S, fnMax, md(ms.x), md(ms.y),
        [mat.Ktc, mat.Krc, mat.Kte, mat.Kre, mat.Kac, mat.Kae, mat.mc != null ? mat.mc : 0, mat.Ci != null ? mat.Ci : 0]]);
    };
    e.setWear = edge => {
      fw(edge);
      const last = wears[wears.length - 1];
      if (!last || last[1] !== edge.Kte || last[2] !== edge.Kre) wears.push([rec.actual.length, edge.Kte, edge.Kre]);
    };
    e.run = n => { fr(n); runs.push([rec.actual.length, n]); };
    const ft = e.takeStats.bind(e), fci = e.chatterIndex.bind(e);
    let took = null, ci = null;
    e.takeStats = () => { const s = ft(); took = ST.map(k => s[k]); return s; };
    e.chatterIndex = () => { const c = fci(); ci = [c.index, c.amp, c.once === undefined ? null : c.once]; return c; };
    const fcol = sim.collect.bind(sim), frames = [];
    sim.collect = (fs, dt) => {
      const VB0 = sim.VB || 0;
      took = null; ci = null;
      fcol(fs, dt);
      const d = sim.display;
      frames.push({ end: rec.actual.length, stats: took, ci, state: d.state, chatterIdx: d.chatterIdx, once: d.once,
        VB0, VB: sim.VB || 0, line: sim.line, lastCutLine: sim.lastCutLine, ap: sim.ap || 0, phiSt: d.phiSt, phiEx: d.phiEx,
        cutTime: fs.cutTime });
    };
    let n = 0;
    while (!sim.done && n < 20000) { sim.step(dtReal); n++; }
    const col = (rows, k) => rows.map(r => r[k]);
    const pack = rows => ({ cx: b64f64(col(rows, 0)), cy: b64f64(col(rows, 1)), tipZ: b64f64(col(rows, 2)),
      R: b64f64(col(rows, 3)), type: col(rows, 4), errVal: b64f64(col(rows, 5)), f: b64f64(col(rows, 6)),
      line: col(rows, 7), vol: b64f64(col(rows, 8)) });
    const P = sim.path, t = sim.activeTool;
    const clr = t.clearanceDeg != null ? t.clearanceDeg : (sim.cfg.tool.clearanceDeg != null ? sim.cfg.tool.clearanceDeg : 10);
    const maps = s => ({ H: z32(Float32Array, s.H), err: z32(Float32Array, s.err), wallD: z32(Float32Array, s.wallD),
      lineCut: z32(Int32Array, s.lineCut), lineWall: z32(Int32Array, s.lineWall), markS: z32(Float32Array, s.markS),
      markF: z32(Float32Array, s.markF), removed: s.removed });
    return { nframes: n, frames, darc: b64f64(darc), nx: sim.stock.nx, ny: sim.stock.ny, hx: sim.stock.hx, hy: sim.stock.hy,
      zBottom: sim.stock.zBottom, toolR: t.R, toolN: t.N, fluteLen: t.fluteLen, type: t.type, clearanceDeg: clr,
      dynBudget: sim.cfg.dynBudget, speed: sim.speed, feedOverride: sim.cfg.feedOverride, spindleOverride: sim.cfg.spindleOverride,
      path: { x: b64f64(P.map(p => p.x)), y: b64f64(P.map(p => p.y)), z: b64f64(P.map(p => p.z)),
        rapid: P.map(p => (p.rapid ? 1 : 0)), line: P.map(p => p.line), F: P.map(p => p.F), S: P.map(p => p.S) },
      cfgs, wears, runs, actual: pack(rec.actual), ideal: pack(rec.ideal), stock: maps(sim.stock), commanded: maps(sim.ideal),
      report: sim.report.pack() };
  });
  console.error('report', (Date.now() - t0) / 1e3, 's');
}

// the parser's moves and the path discretizer (arcs into chords) on short programs
if (cases.paths) {
  const { parseGcode, discretize } = await import(REF + 'gcode.js');
  out.paths = cases.paths.map(({ program, maxSegs }) => {
    const parsed = parseGcode(program);
    const moves = parsed.moves.map(m => ({ from: [m.from.x, m.from.y, m.from.z], to: [m.to.x, m.to.y, m.to.z],
      rapid: m.rapid ? 1 : 0, F: m.F, S: m.S, line: m.line, tool: m.tool === undefined ? null : m.tool,
      arc: m.arc ? { ca: m.arc.ca, cb: m.arc.cb, cw: m.arc.cw ? 1 : 0, axes: m.arc.axes || ['x', 'y', 'z'],
        sweep: m.arc.sweep === undefined ? null : m.arc.sweep } : null }));
    const pts = maxSegs.map(ms => {
      const P = discretize(parsed.moves, ms);
      return { x: b64f64(P.map(p => p.x)), y: b64f64(P.map(p => p.y)), z: b64f64(P.map(p => p.z)),
        rapid: P.map(p => (p.rapid ? 1 : 0)), F: P.map(p => (p.F === undefined ? null : p.F)),
        S: P.map(p => (p.S === undefined ? null : p.S)), line: P.map(p => p.line) };
    });
    return { moves, pts, warnings: parsed.warnings ? parsed.warnings.length : 0 };
  });
  console.error('paths', (Date.now() - t0) / 1e3, 's');
}

// the engine's own call sequence inside whole simulator runs: every configure, setWear,
// setWork, setEngagement, setEndEngagement, run, takeStats and chatterIndex the simulator
// makes, with the maps and the state it set before each run and the engine's state, stats,
// rings and chip events after it
if (cases.englog) {
  const { Simulator } = await import(REF + 'sim.js');
  const { modeSet } = await import(REF + 'frf.js');
  const ST = ['n', 'sumF', 'maxF', 'sumFx', 'sumFy', 'sumFz', 'maxFz', 'sumT', 'maxT', 'sumX', 'sumY', 'minX',
    'maxX', 'minY', 'maxY', 'sumZc', 'wZc', 'sumW', 'maxW', 'sumHb', 'sumB', 'sumRb', 'engTeeth'];
  const rings = e => [e.hist.i, e.hist.count, e.toothSamples.i, e.toothSamples.count, e.wallSamples.i,
    e.wallSamples.count, e.wallSamplesPi.i, e.wallSamplesPi.count];
  out.englog = cases.englog.map(({ program, tool, stock, cell, flex, nonlinear }) => {
    const sim = new Simulator();
    sim.cfg.cell = cell;
    if (flex !== undefined) sim.cfg.workpieceFlex = flex;
    if (nonlinear !== undefined) sim.cfg.cuttingNonlinear = nonlinear;
    Object.assign(sim.cfg.tool, tool);
    sim.load(program, { stock, origin: { x: 0, y: 0, z: 0 } });
    const e = sim.engine;
    const ops = [], Hb = [], Eb = [], pre = [], post = [], ev = [], evn = [], takes = [], cis = [], sl = [], rg = [];
    let inCfg = false;
    const f = {};
    for (const m of ['configure', 'setWear', 'setWork', 'setEngagement', 'setEndEngagement', 'run', 'takeStats', 'chatterIndex']) f[m] = e[m].bind(e);
    e.configure = (t, modal, mat, S, fnMax = 0) => {
      inCfg = true;
      f.configure(t, modal, mat, S, fnMax);
      inCfg = false;
      const ms = modeSet(modal), md = L => L.map(m => [m.k, m.fn, m.zeta]);
      ops.push(['cfg', t.N, t.R, t.helixDeg, t.type, t.fluteLen, t.nonlinear === false ? 0 : 1, t.runout || 0,
        t.runoutPhase || 0, S, fnMax, md(ms.x), md(ms.y),
        [mat.Ktc, mat.Krc, mat.Kte, mat.Kre, mat.Kac, mat.Kae, mat.mc != null ? mat.mc : 0, mat.Ci != null ? mat.Ci : 0],
        [e.stepsPerTooth, e.stepsPerRev, e.decim, e.piStep, e.maxSlices, e.dz, e.dt, e.omega]]);
    };
    e.setWear = edge => { f.setWear(edge); ops.push(['wear', edge.Kte, edge.Kre]); };
    // configure's own setWork(null) is part of configure
    e.setWork = w => { f.setWork(w); if (!inCfg) ops.push(w ? ['work', w.k, w.m, w.zeta, w.nx, w.ny] : ['work']); };
    e.setEngagement = (u, fz, tipZ, ap) => {
      f.setEngagement(u, fz, tipZ, ap);
      ops.push(['eng', u[0], u[1], fz, tipZ, ap, e.tool.rAt ? 1 : 0, e.nz]);
      sl.push(...e.sliceR, ...e.sliceKap, ...e.sliceZ, ...e.sliceB, ...e.sliceLag);
    };
    e.setEndEngagement = (fza, rOut, zAt) => {
      f.setEndEngagement(fza, rOut, zAt);
      ops.push(['end', fza, rOut]);
      rg.push(...e.ringR, ...e.ringCos, ...e.ringZ, ...e.ringDs);
    };
    e.run = n => {
      Hb.push(...e.Hb); Eb.push(...e.Eb);
      pre.push(e.x, e.y, e.vx, e.vy, e.w, e.wv);
      const before = e.chipEvents.length;
      f.run(n);
      const st = e.stats;
      post.push(e.x, e.y, e.vx, e.vy, e.w, e.wv, e.t, e.pr, e.runMeanW, e.stepCount, e.nz, ...rings(e), ...ST.map(k => st[k]));
      const nev = e.chipEvents.length - before;
      evn.push(nev);
      for (let q = before; q < e.chipEvents.length; q++) { const c = e.chipEvents[q]; ev.push(c.vol, c.phi, c.z); }
      ops.push(['run', n]);
    };
    e.takeStats = () => { const s = f.takeStats(); takes.push(...ST.map(k => s[k])); ops.push(['take']); return s; };
    e.chatterIndex = () => { const c = f.chatterIndex(); cis.push(c.index, c.amp, c.once === undefined ? 0 : c.once); ops.push(['ci']); return c; };
    let frames = 0;
    while (!sim.done && frames < 100) { sim.step(1e6); frames++; }
    const R = (b => ({ x: b64f32(b.x), y: b64f32(b.y) }));
    return { ops, Hb: z32(Float32Array, Hb), Eb: z32(Float32Array, Eb), pre: z32(Float64Array, pre),
      post: z32(Float64Array, post), ev: z32(Float64Array, ev), evn, takes: z32(Float64Array, takes),
      cis: b64f64(cis), sl: z32(Float64Array, sl), rg: z32(Float64Array, rg), nst: ST.length,
      final: { hist: z32(Float32Array, [...e.hist.Fx, ...e.hist.Fy, ...e.hist.Fz, ...e.hist.x, ...e.hist.y]),
        ts: R(e.toothSamples), ws: R(e.wallSamples), wsp: R(e.wallSamplesPi) } };
  });
  console.error('englog', (Date.now() - t0) / 1e3, 's');
}

// the engine as Simulator.ensureEngine / step configure it (src/sim.js:252-257,403-404)
if (cases.enginex) {
  out.enginex = cases.enginex.map(c => {
    const e = new CE();
    e.configure({ N: c.tool.N, R: c.tool.R, helixDeg: c.tool.helixDeg, type: c.tool.type, fluteLen: c.tool.fluteLen,
      nonlinear: c.nonlinear, runout: c.tool.runout, runoutPhase: c.tool.runoutPhase }, c.modal, c.mat, c.S, c.fnMax);
    e.setEngagement(c.u, c.fz, c.tipZ, c.ap);
    e.Hb.set(c.Hb.map(v => (v === '-inf' ? -Infinity : v)));
    e.vx = c.v0[0]; e.vy = c.v0[1];
    const Fx = [], Fy = [], x = [], y = [];
    e.takeStats();
    for (let s = 0; s < c.steps; s++) {
      e.run(1);
      const st = e.takeStats();
      Fx.push(st.sumFx); Fy.push(st.sumFy); x.push(e.x); y.push(e.y);
    }
    return { Fx: b64f64(Fx), Fy: b64f64(Fy), x: b64f64(x), y: b64f64(y), nz: e.nz, spt: e.stepsPerTooth };
  });
}

if (cases.tiplut) out.tiplut = cases.tiplut.map(t => b64f32(tipLut(t)));

if (cases.toolModel) out.toolModel = cases.toolModel.map(t => { const m = toolModel(t); return { k: m.k, fn: m.fn, De: m.De, I: m.I, kTool: m.kTool, mEff: m.mEff }; });

if (cases.spindle) {
  const sp = await import(REF + 'spindle.js');
  out.spindle = cases.spindle.map(({ modal, chain, soft, speeds }) => {
    const sf = soft === null ? undefined : soft;
    return { modes: speeds.map(S => sp.modesAt(modal, S, chain, sf)),
      bearing: speeds.map(S => sp.bearingFactor(S, sf)),
      tip: chain ? speeds.map(S => sp.tipFactor(S, chain.kTool, chain.kSp, sf)) : null,
      twoDof: chain && chain.twoDof ? sp.twoDofModes(chain.kTool, chain.twoDof.mTool, chain.kSp, chain.twoDof.mSp, modal.zeta, chain.twoDof.zSp) : null,
      dependent: sp.speedDependent(modal, sf) };
  });
}

// Simulator.lobes (src/sim.js:262-286): the composite curve of a speed-dependent tool tip and
// the diagram at rest, per engagement arc; the modalNow calls each lobes() call makes are
// recorded (speed and the x / y modes) so the nodes' inputs are known exactly
if (cases.clobes) {
  const { Simulator } = await import(REF + 'sim.js');
  const { modeSet } = await import(REF + 'frf.js');
  const { speedDependent } = await import(REF + 'spindle.js');
  const md = L => L.map(m => [m.k, m.fn, m.zeta]);
  const lob = L => L.map(l => ({ rpm: z32(Float64Array, l.map(p => p.rpm)), alim: z32(Float64Array, l.map(p => p.alim)) }));
  out.clobes = cases.clobes.map(({ program, tool, modal, soft, material, arcs }) => {
    const sim = new Simulator();
    if (material) sim.cfg.material = material;
    if (tool) Object.assign(sim.cfg.tool, tool);
    if (modal) Object.assign(sim.cfg.modal, modal);
    if (soft !== undefined) sim.cfg.spindleSoft = soft;
    sim.load(program);
    const orig = sim.modalNow.bind(sim);
    let calls = [];
    sim.modalNow = S => { const m = orig(S); const ms = modeSet(m); calls.push([S, md(ms.x), md(ms.y)]); return m; };
    const runs = arcs.map(([phiSt, phiEx, staticOnly]) => {
      calls = [];
      const L = sim.lobes(phiSt, phiEx, !!staticOnly);
      return { lobes: lob(L), calls };
    });
    return { N: sim.activeTool.N, Ktc: sim.mat.Ktc, Krc: sim.mat.Krc, chain: sim.chain(), modal: sim.cfg.modal,
      soft: sim.cfg.spindleSoft === undefined ? null : sim.cfg.spindleSoft,
      dependent: speedDependent(sim.cfg.modal, sim.cfg.spindleSoft), runs };
  });
  console.error('clobes', (Date.now() - t0) / 1e3, 's');
}

fs.writeFileSync(outPath, JSON.stringify(out));
console.error('done', (Date.now() - t0) / 1e3, 's');
}
main().catch(e => { console.error(e); process.exit(1); });
