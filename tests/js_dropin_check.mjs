// Drives the JavaScript drop-in modules (machinig-sim_amd/js) on the golden cases written
// from the reference and checks they return the reference's answers.
// usage: node js_dropin_check.mjs [lobes|gpu [surrogate model file]]
import fs from 'fs';
import zlib from 'zlib';
import path from 'path';
import { fileURLToPath } from 'url';

const here = path.dirname(fileURLToPath(import.meta.url));
const mod = path.join(here, '..', 'machinig-sim_amd', 'js', 'index.js');
const load = n => JSON.parse(fs.readFileSync(path.join(here, 'golden', 'golden_' + n + '.json'), 'utf8'));
const f64 = s => { const b = Buffer.from(s, 'base64'); return new Float64Array(b.buffer, b.byteOffset, b.length / 8); };
const f32 = s => { const b = Buffer.from(s, 'base64'); return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.length)); };
const i32 = s => { const b = Buffer.from(s, 'base64'); return new Int32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.length)); };
const num = v => (v === 'inf' ? Infinity : v === '-inf' ? -Infinity : v);
let fails = 0;
const check = (ok, what) => { if (!ok) { fails++; console.log('FAIL', what); } };
const sameBits = (a, b) => a.length === b.length && a.every((v, i) => Object.is(v, b[i]));
const zdec = (s, T) => { const b = zlib.inflateSync(Buffer.from(s, 'base64')); return new T(b.buffer.slice(b.byteOffset, b.byteOffset + b.length)); };
const col = (L, k, T = Float64Array) => T.from(L.map(p => p[k]));

// the recorded run's polyline through simSchedule (every sub-step, unbounded budget)
function schedule(m, res) {
  const P = res.path, X = f64(P.x), Y = f64(P.y), Z = f64(P.z);
  const pts = Array.from(X, (x, i) => ({ x, y: Y[i], z: Z[i], rapid: P.rapid[i], F: P.F[i], S: P.S[i], line: P.line[i] }));
  return m.simSchedule(pts, { hx: res.hx, N: res.toolN });
}

function clobesCheck(m, what) {
    // Simulator.lobes with this build's modules: modalNow from modesAt on the recorded modal,
  // chain and softening; the composite curve from compositeLobes, the diagram at rest (and
  // a speed-independent tip) from stabilityLobes of modalNow(0)
  const cl = load('clobes');
  const z64 = s => { const b = zlib.inflateSync(Buffer.from(s, 'base64')); return new Float64Array(b.buffer, b.byteOffset, b.length / 8); };
  cl.cases.forEach((c, k) => {
    const res = cl.results[k], soft = res.soft === null ? undefined : res.soft, md = res.modal;
    const modalNow = S => { const ms = m.modesAt(md, S, res.chain, soft); return { k: md.k, fn: md.fn, zeta: md.zeta, modesX: ms.x, modesY: ms.y }; };
    c.arcs.forEach(([phiSt, phiEx, staticOnly], a) => {
      const base = { N: res.N, Ktc: res.Ktc, Krc: res.Krc, phiSt, phiEx, lobes: 30, rpmMin: 2000, rpmMax: 40000 };
      const dep = !staticOnly && m.speedDependent(md, soft);
      if (dep !== (what === 'gpu')) return;   // the composite curve is device work
      const got = dep ? m.compositeLobes(base, modalNow) : m.stabilityLobes(Object.assign({}, modalNow(0), base));
      const want = res.runs[a].lobes;
      check(got.length === want.length, 'composite lobe count ' + k + '/' + a);
      got.forEach((L, j) => {
        check(sameBits(Float64Array.from(L.map(p => p.rpm)), z64(want[j].rpm)), 'composite rpm ' + k + '/' + a + '/' + j);
        check(sameBits(Float64Array.from(L.map(p => p.alim)), z64(want[j].alim)), 'composite alim ' + k + '/' + a + '/' + j);
      });
    });
  });
}

async function main() {
  const m = await import(mod);
  const what = process.argv[2] || 'gpu';
  if (what === 'lobes') {
    const g = load('lobes');
    g.cases.forEach((c, k) => {
      const got = m.stabilityLobes(c.params), ref = g.results[k];
      check(got.length === ref.lobes.length, 'lobe count ' + k);
      got.forEach((L, j) => {
        check(sameBits(Float64Array.from(L.map(p => p.rpm)), f64(ref.lobes[j].rpm)), 'rpm ' + k + '/' + j);
        check(sameBits(Float64Array.from(L.map(p => p.alim)), f64(ref.lobes[j].alim)), 'alim ' + k + '/' + j);
      });
      c.speeds.forEach((S, i) => check(Object.is(m.alimAt(got, S), num(ref.alimAt[i])), 'alimAt ' + k + '/' + i));
      const d = m.directionalCoeffs(c.params.Krc / c.params.Ktc, c.params.phiSt, c.params.phiEx);
      check(sameBits([d.xx, d.xy, d.yx, d.yy], ref.dc), 'dc ' + k);
    });
    clobesCheck(m, what);
    const tl = load('tiplut');
    tl.cases.forEach((t, k) => check(sameBits(m.tipLut(t), f32(tl.results[k])), 'tipLut ' + k));
    check(m.LUT_DR === 0.02, 'LUT_DR');
    // toolModel and the speed-dependent modes are host code: the reference's doubles
    const tm = load('toolmodel');
    tm.cases.forEach((t, k) => {
      const got = m.toolModel(t), want = tm.results[k];
      check(['k', 'fn', 'De', 'I', 'kTool', 'mEff'].every(f => Object.is(got[f], want[f])), 'toolModel ' + k);
    });
    const sp = load('spindle');
    const same = (a, b) => JSON.stringify(a) === JSON.stringify(b);
    sp.cases.forEach((c, k) => {
      const soft = c.soft === null ? undefined : c.soft, want = sp.results[k];
      check(same(c.speeds.map(S => m.modesAt(c.modal, S, c.chain, soft)), want.modes), 'modesAt ' + k);
      check(same(c.speeds.map(S => m.bearingFactor(S, soft)), want.bearing), 'bearingFactor ' + k);
      if (c.chain) check(same(c.speeds.map(S => m.tipFactor(S, c.chain.kTool, c.chain.kSp, soft)), want.tip), 'tipFactor ' + k);
      if (want.twoDof) check(same(m.twoDofModes(c.chain.kTool, c.chain.twoDof.mTool, c.chain.kSp, c.chain.twoDof.mSp, c.modal.zeta, c.chain.twoDof.zSp), want.twoDof), 'twoDofModes ' + k);
      check(m.speedDependent(c.modal, soft) === want.dependent, 'speedDependent ' + k);
    });
    // immersion is host code over engine.Hb: the reference's answers from its own Hb
    const eg = load('engage');
    eg.cases.forEach((c, k) => {
      const Hb = f32(eg.results[k].Hb);
      c.queries.forEach((q, i) => {
        const me = { engine: { Hb: Hb.subarray(180 * i, 180 * i + 180) } };
        const im = m.immersion.call(me, q.tipZ), want = eg.results[k].imm[i];
        check(want === null ? im === null : im !== null && im.st === want[0] && im.ex === want[1], 'immersion ' + k + '/' + i);
      });
    });
    // the path format and the sub-step schedule (host C++ behind the addon): discretize on the
    // reference parser's moves, simSchedule on the recorded simulator runs' polylines
    const gp = load('paths');
    const obj = v => ({ x: v[0], y: v[1], z: v[2] });
    gp.cases.forEach((c, k) => {
      const moves = gp.results[k].moves.map(mv => Object.assign({}, mv, { from: obj(mv.from), to: obj(mv.to) }));
      c.maxSegs.forEach((ms, j) => {
        const pts = m.discretize(moves, ms), ref = gp.results[k].pts[j];
        check(sameBits(col(pts, 'x'), f64(ref.x)) && sameBits(col(pts, 'y'), f64(ref.y)) && sameBits(col(pts, 'z'), f64(ref.z)) &&
          pts.every((p, i) => p.rapid === !!ref.rapid[i] && p.line === ref.line[i] && p.F === ref.F[i] && p.S === ref.S[i]),
        'discretize ' + k + '/' + j);
      });
    });
    const sx = load('simx');
    sx.results.forEach((res, k) => {
      const sub = m.readSubsteps(schedule(m, res).sub), ide = res.ideal;
      check(sub.length === ide.type.length && sameBits(col(sub, 'x'), f64(ide.cx)) && sameBits(col(sub, 'y'), f64(ide.cy)) &&
        sameBits(col(sub, 'z'), f64(ide.tipZ)) && sub.every((q, i) => q.line === ide.line[i]), 'simSchedule ' + k);
    });
  } else {
    console.log('device', m.native().deviceInfo());
    clobesCheck(m, what);
    const ch = load('chatter');
    ch.cases.forEach((c, k) => {
      const x = c.cell;
      const v = m.chatterAt(x.tool, x.modal, x.mat, x.S, x.arc, x.fz, c.ap, { revs: x.revs, nonlinear: x.nonlinear });
      check(v === (ch.results[k].chatter === 1), 'chatterAt ' + k);
    });
    const td = load('tdl');
    const byOpts = td.cases.map(c => ({ tool: c.cell.tool, modal: c.cell.modal, mat: c.cell.mat, S: c.cell.S, arc: c.cell.arc, fz: c.cell.fz }));
    // one launch per (revs, nonlinear) group, as timeDomainLimits takes shared options
    const groups = {};
    td.cases.forEach((c, k) => { const key = c.cell.revs + '|' + c.cell.nonlinear; (groups[key] = groups[key] || []).push(k); });
    for (const key of Object.keys(groups)) {
      const ks = groups[key], c0 = td.cases[ks[0]];
      const res = m.timeDomainLimits(ks.map(k => byOpts[k]), Object.assign({}, c0.opts, { revs: c0.cell.revs, nonlinear: c0.cell.nonlinear }));
      res.forEach((r, i) => {
        const ref = td.results[ks[i]];
        check(r.alim === ref.alim && r.bracket[0] === ref.lo && r.bracket[1] === num(ref.hi), 'timeDomainLimit ' + ks[i]);
      });
    }
    const one = td.cases[0];
    const r1 = m.timeDomainLimit(one.cell.tool, one.cell.modal, one.cell.mat, one.cell.S, one.cell.arc, one.cell.fz,
      Object.assign({}, one.opts, { revs: one.cell.revs, nonlinear: one.cell.nonlinear }));
    check(r1.alim === td.results[0].alim, 'timeDomainLimit single');
    const ra = await m.timeDomainLimitsAsync([byOpts[0]], Object.assign({}, one.opts, { revs: one.cell.revs, nonlinear: one.cell.nonlinear }));
    check(ra[0].alim === td.results[0].alim, 'timeDomainLimitsAsync');
    const st = load('stock');
    st.cases.forEach((c, k) => {
      const s = new m.Stock(c.w, c.l, c.h, c.cell), ref = st.results[k];
      if (c.wallBand) s.wallBand = c.wallBand;
      const luts = (c.tools || []).map(m.tipLut);
      luts.forEach((L, i) => check(sameBits(L, f32(ref.luts[i])), 'stock tipLut ' + k + '/' + i));
      const args = p => [p.cx, p.cy, p.tipZ, p.R, p.type, p.errVal, p.mark || null, p.lut === undefined ? null : luts[p.lut]];
      const ent = p => ({ cx: p.cx, cy: p.cy, tipZ: p.tipZ, R: p.R, type: p.type, errVal: p.errVal, line: p.line,
        mark: p.mark || null, lut: p.lut === undefined ? null : luts[p.lut] });
      const half = c.clearAt === undefined ? c.stamps.length >> 1 : c.clearAt;
      const vol = [], dirty = [], disk = [];
      const ask = () => (c.queries || []).forEach(([x, y, r, a]) => { const q = s.maxInDisk(x, y, r, num(a)); disk.push([q.h, q.x, q.y]); });
      // the first part one call at a time, the rest as one batch
      for (const p of c.stamps.slice(0, half)) { s.curLine = p.line; vol.push(s.stamp(...args(p))); }
      if (c.clearAt !== undefined) { const d = s.dirty; dirty.push([d.i0, d.j0, d.i1, d.j1]); ask(); s.clearDirty(); }
      vol.push(...s.stampBatch(c.stamps.slice(half).map(ent)));
      check(sameBits(s.markS, f32(ref.markS)) && sameBits(s.markF, f32(ref.markF)), 'marks ' + k);
      if (c.queries) {
        const d = s.dirty; dirty.push([d.i0, d.j0, d.i1, d.j1]); ask();
        check(JSON.stringify(dirty) === JSON.stringify(ref.dirty.map(r => r.map(num))), 'dirty ' + k);
        check(disk.length === ref.disk.length && disk.every((q, i) => Object.is(q[0], num(ref.disk[i][0])) &&
          q[1] === ref.disk[i][1] && q[2] === ref.disk[i][2]), 'maxInDisk ' + k);
        const bm = [];
        for (let bj = 0; bj < s.bny; bj++) for (let bi = 0; bi < s.bnx; bi++) bm.push(s.blockMax(bi, bj));
        check(sameBits(Float32Array.from(bm), f32(ref.bmax)), 'blockMax ' + k);
        s.clearDirty();
        for (const w of c.walls) s.markWall(...w);
        const d2 = s.dirty;
        check(sameBits(s.wallErr, f32(ref.wallErr)), 'markWall ' + k);
        check(JSON.stringify([d2.i0, d2.j0, d2.i1, d2.j1]) === JSON.stringify(ref.wallDirty.map(num)), 'markWall dirty ' + k);
      }
      check(sameBits(s.H, f32(ref.H)), 'H ' + k);
      check(sameBits(s.err, f32(ref.err)) && sameBits(s.wallD, f32(ref.wallD)), 'err/wallD ' + k);
      check(sameBits(s.lineCut, i32(ref.lineCut)) && sameBits(s.lineWall, i32(ref.lineWall)), 'lines ' + k);
      check(vol.every((v, i) => Math.abs(v - ref.vol[i]) <= 1e-12 * Math.max(1, Math.abs(ref.vol[i]))), 'vol ' + k);
      check(Math.abs(s.removed - ref.removed) <= 1e-9 * Math.max(1, ref.removed), 'removed ' + k);
      c.samples.forEach(([x, y], i) => check(Object.is(s.sample(x, y), num(ref.sample[i])), 'sample ' + k + '/' + i));
      s.dispose();
    });
    // the simulator's engagement sampling as methods of a stand-in simulator on this Stock
    const eg = load('engage');
    eg.cases.forEach((c, k) => {
      const sc = st.cases[c.stock], ref = eg.results[k];
      const s = new m.Stock(sc.w, sc.l, sc.h, sc.cell);
      if (sc.wallBand) s.wallBand = sc.wallBand;
      const luts = (sc.tools || []).map(m.tipLut);
      s.stampBatch(sc.stamps.map(p => ({ cx: p.cx, cy: p.cy, tipZ: p.tipZ, R: p.R, type: p.type, errVal: p.errVal,
        line: p.line, mark: p.mark || null, lut: p.lut === undefined ? null : luts[p.lut] })));
      const HbRef = f32(ref.Hb);
      c.queries.forEach((q, i) => {
        const me = { engine: { Hb: new Float32Array(180) }, activeTool: { R: q.R }, stock: s };
        const h = m.fillEngagement.call(me, q.cx, q.cy, [q.ux, q.uy]);
        check(Object.is(h, num(ref.hmax[i])) && sameBits(me.engine.Hb, HbRef.subarray(180 * i, 180 * i + 180)), 'fillEngagement ' + k + '/' + i);
        const im = m.immersion.call(me, q.tipZ), want = ref.imm[i];
        check(want === null ? im === null : im !== null && im.st === want[0] && im.ex === want[1], 'immersion ' + k + '/' + i);
      });
      const b = m.engagementBatch(s, c.queries);
      check(b.every((r, i) => Object.is(r.hmax, num(ref.hmax[i]))), 'engagementBatch ' + k);
      s.dispose();
    });
    // the CuttingEngine class over the simulator's recorded call sequences (linear force law,
    // flat and ball tools): state, stats, chip events, takeStats, chatterIndex and rings
    const eq = (a, b) => a === b || (a !== a && b !== b);
    const STATS = ['n', 'sumF', 'maxF', 'sumFx', 'sumFy', 'sumFz', 'maxFz', 'sumT', 'maxT', 'sumX', 'sumY', 'minX',
      'maxX', 'minY', 'maxY', 'sumZc', 'wZc', 'sumW', 'maxW', 'sumHb', 'sumB', 'sumRb', 'engTeeth'];
    const el = load('englog');
    let replayed = 0;
    el.cases.forEach((c, k) => {
      if (c.nonlinear !== false || c.tool.type === 'bull') return;
      const res = el.results[k];
      const Hb = zdec(res.Hb, Float32Array), Eb = zdec(res.Eb, Float32Array), pre = zdec(res.pre, Float64Array);
      const post = zdec(res.post, Float64Array), ev = zdec(res.ev, Float64Array), takes = zdec(res.takes, Float64Array);
      const cis = f64(res.cis), K = 19 + res.nst;
      const e = new m.CuttingEngine();
      let r = 0, evo = 0, it = 0, ic = 0, bad = 0, tool = null;
      for (const op of res.ops) {
        if (op[0] === 'cfg') {
          const [, N, R, hd, type, fl, nl, ro, rph, S, fnMax, mx, my, mt] = op;
          tool = { N, R, helixDeg: hd, type, fluteLen: fl, nonlinear: nl === 1, runout: ro, runoutPhase: rph };
          const modes = L => L.map(q => ({ k: q[0], fn: q[1], zeta: q[2] }));
          e.configure(tool, { k: mx[0][0], fn: mx[0][1], zeta: mx[0][2], modesX: modes(mx), modesY: modes(my) },
            { Ktc: mt[0], Krc: mt[1], Kte: mt[2], Kre: mt[3], Kac: mt[4], Kae: mt[5], mc: mt[6], Ci: mt[7] }, S, fnMax);
          if (e.stepsPerTooth !== op[14][0] || e.decim !== op[14][2] || e.dt !== op[14][6]) bad++;
        } else if (op[0] === 'wear') e.setWear({ Kte: op[1], Kre: op[2] });
        else if (op[0] === 'work') e.setWork(op.length > 1 ? { k: op[1], m: op[2], zeta: op[3], nx: op[4], ny: op[5] } : null);
        else if (op[0] === 'eng') { e.setEngagement([op[1], op[2]], op[3], op[4], op[5]); if (e.nz !== op[7]) bad++; }
        else if (op[0] === 'end') {
          const R = tool.R;
          e.setEndEngagement(op[1], op[2], tool.type === 'ball' ? (q => R - Math.sqrt(Math.max(0, R * R - q * q))) : undefined);
        } else if (op[0] === 'run') {
          e.Hb.set(Hb.subarray(180 * r, 180 * r + 180)); e.Eb.set(Eb.subarray(216 * r, 216 * r + 216));
          [e.x, e.y, e.vx, e.vy, e.w, e.wv] = pre.subarray(6 * r, 6 * r + 6);
          const before = e.chipEvents.length;
          e.run(op[1]);
          const p = post.subarray(K * r, K * r + K);
          const got = [e.x, e.y, e.vx, e.vy, e.w, e.wv, e.t, e.pr, e.runMeanW, e.stepCount, e.nz, e.hist.i, e.hist.count,
            e.toothSamples.i, e.toothSamples.count, e.wallSamples.i, e.wallSamples.count, e.wallSamplesPi.i, e.wallSamplesPi.count];
          if (!got.every((v, i) => eq(v, p[i]))) bad++;
          if (!STATS.every((f, i) => eq(e.stats[f], p[19 + i]))) bad++;
          const nev = res.evn[r], mine = e.chipEvents.slice(before);
          if (mine.length !== nev || !mine.every((q, i) => eq(q.vol, ev[3 * (evo + i)]) && eq(q.phi, ev[3 * (evo + i) + 1]) &&
            eq(q.z, ev[3 * (evo + i) + 2]) && q.ux === e.ux && q.vy === e.vy_)) bad++;
          evo += nev;
          r++;
        } else if (op[0] === 'take') {
          const s = e.takeStats();
          if (!STATS.every((f, i) => eq(s[f], takes[res.nst * it + i]))) bad++;
          it++;
        } else if (op[0] === 'ci') {
          const ci = e.chatterIndex();
          if (!(eq(ci.index, cis[3 * ic]) && eq(ci.amp, cis[3 * ic + 1]) && eq(ci.once === undefined ? 0 : ci.once, cis[3 * ic + 2]))) bad++;
          ic++;
        }
        if (bad) break;
      }
      const fin = res.final, hist = zdec(fin.hist, Float32Array);
      if (!bad) {
        const H = 4096;
        ['Fx', 'Fy', 'Fz', 'x', 'y'].forEach((f, j) => { if (!sameBits(e.hist[f], hist.subarray(j * H, j * H + H))) bad++; });
        [['ts', e.toothSamples], ['ws', e.wallSamples], ['wsp', e.wallSamplesPi]].forEach(([n, b]) => {
          if (!sameBits(b.x, f32(fin[n].x)) || !sameBits(b.y, f32(fin[n].y))) bad++;
        });
      }
      check(bad === 0 && r > 100, 'CuttingEngine replay ' + k + ' (' + c.tool.type + ', ' + r + ' runs)');
      replayed++;
    });
    check(replayed === 3, 'CuttingEngine replays');
    const { handle } = await import(path.join(here, '..', 'machinig-sim_amd', 'js', 'stabworker.js'));
    const msgs = [];
    const c0 = td.cases[0].cell;
    await handle({ key: 7, points: [{ S: c0.S, modal: c0.modal }], tool: c0.tool, mat: c0.mat, arc: c0.arc, fz: c0.fz, nonlinear: c0.nonlinear }, r => msgs.push(r));
    check(msgs.length === 2 && msgs[1].done === true && msgs[0].key === 7, 'stabworker protocol');
    // the surrogate-guided search gives the plain search's answers whatever the prior
    if (process.argv[3]) {
      const sg = m.Surrogate.load(process.argv[3]);
      for (const key of Object.keys(groups)) {
        const ks = groups[key], c0 = td.cases[ks[0]];
        const o = Object.assign({}, c0.opts, { revs: c0.cell.revs, nonlinear: c0.cell.nonlinear });
        const res = m.timeDomainLimitsGuided(ks.map(k => byOpts[k]), sg, o);
        res.forEach((r, i) => {
          const ref = td.results[ks[i]];
          check(r.alim === ref.alim && r.bracket[0] === ref.lo && r.bracket[1] === num(ref.hi), 'timeDomainLimitsGuided ' + ks[i]);
        });
      }
      check(sg.predict([byOpts[0]], { revs: c0.revs, nonlinear: c0.nonlinear }).every(Number.isFinite), 'surrogate predict');
    }
    // the simulator's sub-step loop through simRun: the recorded rigid-workpiece runs (flat,
    // ball, bull-nose, an imported edge) in one call, their deflected stamps, vibration values, step counts and
    // both stocks' maps as the reference's
    const sx = load('simx'), progs = [];
    sx.cases.forEach((c, k) => {
      const res = sx.results[k], st = c.stock;
      const act = new m.Stock(st.w, st.l, st.h, c.cell), ide = new m.Stock(st.w, st.l, st.h, c.cell);
      ide.wallBand = 2.0;
      const [, N, R, hd, typ, fl, nl, ro, rph, S, fnMax, mx, my, mt] = res.cfgs[0];
      const e = new m.CuttingEngine(), modes = L => L.map(q => ({ k: q[0], fn: q[1], zeta: q[2] }));
      e.configure({ N, R, helixDeg: hd, type: typ, fluteLen: fl, nonlinear: nl === 1, runout: ro, runoutPhase: rph },
        { k: mx[0][0], fn: mx[0][1], zeta: mx[0][2], modesX: modes(mx), modesY: modes(my) },
        { Ktc: mt[0], Krc: mt[1], Kte: mt[2], Kre: mt[3], Kac: mt[4], Kae: mt[5], mc: mt[6], Ci: mt[7] }, S, fnMax);
      e.setWear({ Kte: res.wears[0][1], Kre: res.wears[0][2] });
      const tool = { R: res.toolR, fluteLen: res.fluteLen, type: res.type };
      if (res.type === 'bull') Object.assign(tool, { cornerR: c.tool.cornerR, lut: m.tipLut({ type: 'bull', D: 2 * res.toolR, cornerR: c.tool.cornerR }) });
      if (res.type === 'profile') {
        const edge = { dz: c.tool.edge.dz, r: Float32Array.from(c.tool.edge.r) };
        Object.assign(tool, { edge, lut: m.tipLut({ type: 'profile', D: 2 * res.toolR, edge }) });
      }
      progs.push({ stock: act, ideal: ide, engine: e, schedule: schedule(m, res), tool, share: e.stepsPerTooth / 4 });
    });
    const ran = m.simRun(progs);
    ran.forEach((r, k) => {
      const res = sx.results[k], p = progs[k], sub = m.readSubsteps(p.schedule.sub), a = res.actual;
      const runs = new Map(res.runs.map(([i, n]) => [i, n]));
      check(sameBits(Float64Array.from(sub, (q, i) => q.x + r.out[i].dx), f64(a.cx)) &&
        sameBits(Float64Array.from(sub, (q, i) => q.y + r.out[i].dy), f64(a.cy)), 'simRun stamps ' + k);
      check(sameBits(col(r.out, 'errVal'), f64(a.errVal)), 'simRun errVal ' + k);
      check(r.out.every((o, i) => o.nsteps === (runs.get(i) || 0)), 'simRun steps ' + k);
      check(sameBits(p.stock.H, zdec(res.stock.H, Float32Array)) && sameBits(p.ideal.H, zdec(res.commanded.H, Float32Array)),
        'simRun maps ' + k);
      check(Math.abs(p.stock.removed - res.stock.removed) <= 1e-12 * Math.max(1, res.stock.removed), 'simRun removed ' + k);
      check(p.engine.t > 0 && p.engine.stepCount > 0, 'simRun engine state ' + k);
      p.stock.dispose(); p.ideal.dispose();
    });
  }
  console.log(fails ? fails + ' failures' : 'all ' + what + ' checks passed');
  process.exit(fails ? 1 : 0);
}
main().catch(e => { console.error(e); process.exit(1); });
