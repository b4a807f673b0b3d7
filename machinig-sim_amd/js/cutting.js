// machining-sim-mi355x/cutting.js — drop-in for the reference's CuttingEngine
// (src/cutting.js:21-326) whose state lives on the GPU between calls.
//
// The methods keep the reference's names, arguments and effects: configure, setWear,
// setWork, setEngagement, setEndEngagement, run, chatterIndex, takeStats.  The public
// fields a caller writes (Hb, Eb, x, y, vx, vy, w, wv, decim, hist.i) are host values sent
// with the next run; the ones it reads (x, y, vx, vy, w, wv, t, runMeanW, wk, wnx, wny, nz,
// stepsPerTooth, stepsPerRev, dt, omega, stats, chipEvents, hist, toothSamples,
// wallSamples, wallSamplesPi) are refreshed after every run.  The sample rings are copied
// from the device only when read.
import { native } from './native.js';
import { writeCell, CELL_BYTES } from './pack.js';

export const BINS = 180;
export const END_BINS = 36;
export const END_RINGS = 6;
export const H_REF = 0.05;
const MAX_SLICES = 64, HIST_N = 4096, RING_N = 64;

const STATS = ['n', 'sumF', 'maxF', 'sumFx', 'sumFy', 'sumFz', 'maxFz', 'sumT', 'maxT', 'sumX', 'sumY', 'minX',
  'maxX', 'minY', 'maxY', 'sumZc', 'wZc', 'sumW', 'maxW', 'sumHb', 'sumB', 'sumRb', 'engTeeth'];

function newStats() {
  return { n: 0, sumF: 0, maxF: 0, sumFx: 0, sumFy: 0, sumFz: 0, maxFz: 0, sumT: 0, maxT: 0, sumX: 0, sumY: 0,
    minX: Infinity, maxX: -Infinity, minY: Infinity, maxY: -Infinity, sumZc: 0, wZc: 0, sumW: 0, maxW: 0, sumHb: 0,
    sumB: 0, sumRb: 0, engTeeth: 0 };
}

// a ring whose arrays are copied from the device on first read after a run
function lazyRing(eng, n, key) {
  const r = { n, i: 0, count: 0 };
  Object.defineProperty(r, 'x', { enumerable: true, get: () => eng._rings()[key].x });
  Object.defineProperty(r, 'y', { enumerable: true, get: () => eng._rings()[key].y });
  return r;
}

export class CuttingEngine {
  constructor() {
    this._h = native().engineCreate();
    this.Hb = new Float32Array(BINS);
    this.Eb = new Float32Array(END_BINS * END_RINGS);
    this.fza = 0; this.endR = 0;
    this.stats = newStats();
    const hist = { n: HIST_N, i: 0, count: 0 };
    for (const k of ['Fx', 'Fy', 'Fz', 'x', 'y'])
      Object.defineProperty(hist, k, { enumerable: true, get: () => this._rings().hist[k] });
    this.hist = hist;
    this.toothSamples = lazyRing(this, RING_N, 'ts');
    this.wallSamples = lazyRing(this, RING_N, 'ws');
    this.wallSamplesPi = lazyRing(this, RING_N, 'wsp');
    this.chipEvents = [];
    this._ringCache = null;
    this._cell = new ArrayBuffer(CELL_BYTES);
  }

  _rings() {
    if (!this._ringCache) {
      const f = native().engineRings(this._h);
      const H = HIST_N, R = RING_N, o = 5 * H;
      this._ringCache = {
        hist: { Fx: f.subarray(0, H), Fy: f.subarray(H, 2 * H), Fz: f.subarray(2 * H, 3 * H), x: f.subarray(3 * H, 4 * H), y: f.subarray(4 * H, 5 * H) },
        ts: { x: f.subarray(o, o + R), y: f.subarray(o + R, o + 2 * R) },
        ws: { x: f.subarray(o + 2 * R, o + 3 * R), y: f.subarray(o + 3 * R, o + 4 * R) },
        wsp: { x: f.subarray(o + 4 * R, o + 5 * R), y: f.subarray(o + 5 * R, o + 6 * R) },
      };
    }
    return this._ringCache;
  }

  // configure(tool, modal, mat, S, fnMax), src/cutting.js:39-82
  configure(tool, modal, mat, S, fnMax = 0) {
    this.tool = tool; this.mat = mat; this.modal = modal;
    writeCell(new DataView(this._cell), 0, tool, modal, mat, S, [0, 0], 0, 1, tool.nonlinear !== false);
    const info = native().engineConfigure(this._h, this._cell, fnMax);
    this.N = tool.N; this.R = tool.R; this.tanB = Math.tan(tool.helixDeg * Math.PI / 180);
    const e = tool.runout || 0, g0 = tool.runoutPhase || 0;
    this.dR = Float64Array.from({ length: tool.N }, (_, i) => e * Math.cos(2 * Math.PI * i / tool.N - g0));
    this.dt = info[0]; this.omega = info[1]; this.Ttooth = info[2]; this.dz = info[3]; this.S = info[4];
    this.stepsPerTooth = info[5]; this.stepsPerRev = info[6]; this.decim = info[7]; this.piStep = info[8];
    this.maxSlices = info[9];
    this.x0 = 0; this.y0 = 0; this.vx0 = 0; this.vy0 = 0;
    this.pr = 0;
    this.x = 0; this.y = 0; this.vx = 0; this.vy = 0; this.t = 0;
    this.runMeanW = 0; this.wk = 0; this.wm = 1; this.wc = 0; this.w = 0; this.wv = 0; this.wnx = 1; this.wny = 0;
    this.fza = 0;
    this.toothEngaged = new Uint8Array(this.N);
    for (const b of [this.wallSamples, this.wallSamplesPi]) { b.i = 0; b.count = 0; }
    this.stepCount = 0;
    this._ringCache = null;
  }

  // setWear(edge), src/cutting.js:85: used by the next runs, the material's where undefined
  setWear(edge) { this.Kte = edge.Kte; this.Kre = edge.Kre; }

  // setWork(wall), src/cutting.js:88-96
  setWork(wall) {
    native().engineSetWork(this._h, wall ? Float64Array.of(wall.k, wall.m, wall.zeta, wall.nx, wall.ny) : null);
    if (!wall) { this.runMeanW = 0; this.wk = 0; this.wm = 1; this.wc = 0; this.w = 0; this.wv = 0; this.wnx = 1; this.wny = 0; return; }
    if (this.wk && wall.nx * this.wnx + wall.ny * this.wny < 0.8) { this.w = 0; this.wv = 0; }
    this.wk = wall.k * 1e3; this.wm = wall.m; this.wc = 2 * wall.zeta * Math.sqrt(this.wk * wall.m);
    this.wnx = wall.nx; this.wny = wall.ny;
  }

  // setEngagement(u, fz, tipZ, ap), src/cutting.js:98-124; a tool profile (tool.rAt) is
  // evaluated here, on the host, and its slice radii sent with the engagement
  setEngagement(u, fz, tipZ, ap) {
    this.ux = u[0]; this.uy = u[1]; this.vx_ = -u[1]; this.vy_ = u[0];
    this.fz = fz; this.tipZ = tipZ;
    const nz = ap <= 0 ? 0 : Math.min(this.maxSlices, Math.ceil(ap / this.dz));
    this.nz = nz;
    this.db = this.dz;
    this.sliceB = Float64Array.from({ length: nz }, (_, k) => Math.min(this.dz, ap - k * this.dz));
    this.sliceZ = new Float64Array(nz);
    this.sliceLag = new Float64Array(nz);
    this.sliceR = new Float64Array(nz);
    this.sliceKap = new Float64Array(nz);
    const R = this.R, rAt = this.tool.rAt;
    for (let k = 0; k < nz; k++) {
      const zr = k * this.dz + 0.5 * this.sliceB[k];
      this.sliceZ[k] = tipZ + zr;
      let r = R, kap = 1;
      if (rAt) {
        r = Math.max(1e-3, rAt(zr));
        const dr = (rAt(zr + 0.05) - rAt(Math.max(0, zr - 0.05))) / (zr < 0.05 ? 0.05 + zr : 0.1);
        kap = 1 / Math.sqrt(1 + dr * dr);
      } else if (this.tool.type === 'ball' && zr < R) { r = Math.sqrt(Math.max(1e-6, R * R - (R - zr) ** 2)); kap = r / R; }
      this.sliceR[k] = r; this.sliceKap[k] = kap;
      this.sliceLag[k] = zr * this.tanB / R;
    }
    let sr = null, sk = null;
    if (rAt) {
      sr = new Float64Array(MAX_SLICES); sk = new Float64Array(MAX_SLICES);
      sr.set(this.sliceR); sk.set(this.sliceKap);
    }
    native().engineSetEngagement(this._h, u[0], u[1], fz, tipZ, ap, sr, sk);
  }

  // setEndEngagement(fza, rOut, zAt), src/cutting.js:130-143
  setEndEngagement(fza, rOut, zAt = () => 0) {
    this.fza = fza > 0 ? fza : 0; this.endR = rOut;
    const dr = rOut / END_RINGS;
    this.ringR = new Float64Array(END_RINGS); this.ringCos = new Float64Array(END_RINGS); this.ringZ = new Float64Array(END_RINGS);
    this.ringDs = new Float64Array(END_RINGS);
    for (let j = 0; j < END_RINGS; j++) {
      const r = (j + 0.5) * dr;
      const slope = (zAt(Math.min(rOut, r + 0.25 * dr)) - zAt(Math.max(0, r - 0.25 * dr))) / (0.5 * dr);
      const c = Math.max(0.22, 1 / Math.sqrt(1 + slope * slope));
      this.ringR[j] = r; this.ringCos[j] = c; this.ringDs[j] = dr / c;
      this.ringZ[j] = zAt(r);
    }
    const g = new Float64Array(4 * END_RINGS);
    g.set(this.ringR, 0); g.set(this.ringCos, END_RINGS); g.set(this.ringZ, 2 * END_RINGS); g.set(this.ringDs, 3 * END_RINGS);
    native().engineSetEndEngagement(this._h, fza, rOut, g);
  }

  // run(nSteps), src/cutting.js:146-296, on the device
  run(nSteps) {
    const mat = this.mat;
    const Kte = this.Kte != null ? this.Kte : mat.Kte, Kre = this.Kre != null ? this.Kre : mat.Kre;
    const set = Float64Array.of(this.x, this.y, this.vx, this.vy, this.w, this.wv, this.t, this.decim, this.hist.i);
    this._sync(native().engineRun(this._h, nSteps, this.Hb, this.Eb, set, Kte, Kre), true);
  }

  // the host fields from the device's state vector (engineRun / engineState); events: the
  // run's chip events are appended with the engagement's directions
  _sync(v, events) {
    this.x = this.x0 = v[0]; this.y = this.y0 = v[1]; this.vx = this.vx0 = v[2]; this.vy = this.vy0 = v[3];
    this.w = v[4]; this.wv = v[5]; this.t = v[6]; this.runMeanW = v[7];
    this.stepCount = v[11]; this.pr = v[12];
    this.hist.i = v[15]; this.hist.count = v[16];
    this.toothSamples.i = v[17]; this.toothSamples.count = v[18];
    this.wallSamples.i = v[19]; this.wallSamples.count = v[20];
    this.wallSamplesPi.i = v[21]; this.wallSamplesPi.count = v[22];
    const st = this.stats;
    for (let k = 0; k < STATS.length; k++) st[STATS[k]] = v[23 + k];
    this._ringCache = null;
    const nev = v[46];
    if (events && nev > 0) {
      const ev = native().engineChipEvents(this._h);
      const ux = this.ux, uy = this.uy, vx = this.vx_, vy = this.vy_;
      for (let q = 0; q < nev; q++) this.chipEvents.push({ vol: ev[3 * q], phi: ev[3 * q + 1], z: ev[3 * q + 2], ux, uy, vx, vy });
    }
  }

  // chatterIndex(), src/cutting.js:299-318
  chatterIndex() {
    if (Math.min(this.toothSamples.count, 24) < 8) return { index: 0, amp: 0 };
    const c = native().engineChatterIndex(this._h);
    return { index: c[0], amp: c[1], once: c[2] };
  }

  takeStats() { const s = this.stats; native().engineTakeStats(this._h); this.stats = newStats(); return s; }
}
