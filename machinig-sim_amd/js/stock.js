// Z-heightfield stock with the reference's interface (src/stock.js:8-165), kept resident
// in MI355X HBM.  stamp() / stampBatch() remove material on the device with the
// reference's sequential semantics (H, err, wallD, lineCut, lineWall, markS, markF and the
// dirty rectangle bit-identical); blockMax() and maxInDisk() answer from the device as the
// reference's block walk does.  The public typed arrays are host mirrors refreshed from
// the device on first access after a change.  A caller that writes H directly calls
// uploadH() afterwards.
import { native } from './native.js';
import { packStamps, packStampOpts } from './pack.js';

export const LUT_DR = 0.02;   // radial step of the tip-shape table [mm] (src/assembly.js:8)
const TIP = { flat: 0, ball: 1, bull: 2, edge: 3 };

// tipLut(t) (src/assembly.js:38-54): height of the cutter surface above the tip by radius
export function tipLut(t) {
  const e = t.edge;
  const shape = e ? TIP.edge : t.type === 'ball' ? TIP.ball : t.type === 'bull' ? TIP.bull : TIP.flat;
  return native().tipLut(shape, t.D, t.cornerR || 0, e ? Float64Array.from(e.r) : null, e ? e.dz : 0);
}

export class Stock {
  constructor(w, l, h, cell) {
    const c = cell === undefined ? 0.3 : cell;
    this.x0 = 0; this.y0 = 0; this.w = w; this.l = l;
    this.zTop = 0; this.zBottom = -h;
    this.handle = native().stockCreate(w, l, h, c);
    const d = native().stockDims(this.handle);
    this.nx = d.nx; this.ny = d.ny; this.hx = d.hx; this.hy = d.hy;
    this.B = 8;
    this.bnx = Math.ceil(this.nx / this.B); this.bny = Math.ceil(this.ny / this.B);
    this.curLine = 0;
    this.removed = 0;
    this._wallBand = 0;
    this._mirror = null;
    this._bmax = null;
    // wall errors are written by markWall only and read by the viewers: a host array
    this.wallErr = new Float32Array(this.nx * this.ny);
    this.clearDirty();
  }

  get wallBand() { return this._wallBand; }
  set wallBand(v) { this._wallBand = v || 0; native().stockSetWallBand(this.handle, this._wallBand); }

  _sync() {
    if (this._mirror) return this._mirror;
    const n = this.nx * this.ny;
    const m = { H: new Float32Array(n), err: new Float32Array(n), wallD: new Float32Array(n),
      lineCut: new Int32Array(n), lineWall: new Int32Array(n), markS: new Float32Array(n), markF: new Float32Array(n) };
    native().stockRead(this.handle, m.H, m.err, m.wallD, m.lineCut, m.lineWall);
    native().stockReadMarks(this.handle, m.markS, m.markF);
    this._mirror = m;
    return m;
  }
  get H() { return this._sync().H; }
  get err() { return this._sync().err; }
  get wallD() { return this._sync().wallD; }
  get lineCut() { return this._sync().lineCut; }
  get lineWall() { return this._sync().lineWall; }
  get markS() { return this._sync().markS; }
  get markF() { return this._sync().markF; }
  _changed() { this._mirror = null; this._bmax = null; }

  // push a host-edited H to the device
  uploadH(H) { native().stockWriteH(this.handle, H || this._sync().H); this._changed(); }

  // this.dirty {i0, j0, i1, j1} (Infinity / -1 when empty) is a plain object as in the
  // reference; each batch folds in the boxes of its stamps that changed a node
  clearDirty() { this.dirty = { i0: Infinity, j0: Infinity, i1: -1, j1: -1 }; }
  _foldDirty() {
    const r = native().stockDirty(this.handle), d = this.dirty;
    if (r[2] < 0) return;
    native().stockClearDirty(this.handle);
    d.i0 = Math.min(d.i0, r[0]); d.j0 = Math.min(d.j0, r[1]); d.i1 = Math.max(d.i1, r[2]); d.j1 = Math.max(d.j1, r[3]);
  }
  // block maxima are kept exact on the device; marking blocks is implied by every change
  markBlocks() {}
  blockMax(bi, bj) {
    if (!this._bmax) this._bmax = native().stockBlockMax(this.handle);
    return this._bmax[bj * this.bnx + bi];
  }
  // highest material inside the circle (cx, cy, r) and where it is; -Infinity if none
  maxInDisk(cx, cy, r, above = -Infinity) {
    const o = this.maxInDiskBatch([[cx, cy, r, above]]);
    return o[0];
  }
  // many queries in one launch: [[cx, cy, r, above?], ...] -> [{h, x, y}, ...]
  maxInDiskBatch(list) {
    const q = new Float64Array(4 * list.length);
    list.forEach((a, i) => { q[4 * i] = a[0]; q[4 * i + 1] = a[1]; q[4 * i + 2] = a[2]; q[4 * i + 3] = a.length > 3 ? a[3] : -Infinity; });
    const o = native().stockMaxInDisk(this.handle, q);
    return list.map((_, i) => ({ h: o[3 * i], x: o[3 * i + 1], y: o[3 * i + 2] }));
  }
  inside(x, y) { return x >= 0 && y >= 0 && x <= this.w && y <= this.l; }

  // Stock.markWall (src/stock.js:101-116): val on the wall cells just outside the cutter
  // (R < r <= R + two cells) on the side facing (nx, ny); its box joins the dirty rectangle
  markWall(cx, cy, R, nx, ny, val) {
    const Ro = R + 2 * Math.max(this.hx, this.hy), hx = this.hx, hy = this.hy, W = this.wallErr;
    const i0 = Math.max(0, Math.floor((cx - Ro) / hx)), i1 = Math.min(this.nx - 1, Math.ceil((cx + Ro) / hx));
    const j0 = Math.max(0, Math.floor((cy - Ro) / hy)), j1 = Math.min(this.ny - 1, Math.ceil((cy + Ro) / hy));
    for (let j = j0; j <= j1; j++) {
      const dy = j * hy - cy;
      for (let i = i0; i <= i1; i++) {
        const dx = i * hx - cx, r = Math.hypot(dx, dy);
        const outside = !(r <= R || r > Ro), facing = !((dx * nx + dy * ny) < 0.5 * r);
        if (outside && facing) W[j * this.nx + i] = val;
      }
    }
    if (i0 <= i1 && j0 <= j1) {
      const d = this.dirty;
      d.i0 = Math.min(d.i0, i0); d.j0 = Math.min(d.j0, j0); d.i1 = Math.max(d.i1, i1); d.j1 = Math.max(d.j1, j1);
    }
  }

  // remove material under the tool ('flat' | 'ball'; mark {ux, uy, f} for the floor's feed
  // marks, lut a tipLut table); returns the removed volume [mm^3]
  stamp(cx, cy, tipZ, R, type, errVal, mark = null, lut = null) {
    return this.stampBatch([{ cx, cy, tipZ, R, type, errVal, line: this.curLine, mark, lut }])[0];
  }

  // many stamps in one launch, applied in order: [{cx, cy, tipZ, R, type, errVal, line, mark?, lut?}]
  stampBatch(list) {
    if (!list.length) return new Float64Array(0);
    const o = packStampOpts(list);
    const vol = o ? native().stampBatch(this.handle, packStamps(list), o.opts, o.luts, o.lutOff, o.lutLen)
      : native().stampBatch(this.handle, packStamps(list));
    for (let i = 0; i < list.length; i++) this.removed += vol[i];
    this._changed();
    this._foldDirty();
    return vol;
  }

  // bilinear height; outside the stock -Infinity
  sample(x, y) { return this.sampleBatch(new Float64Array([x, y]))[0]; }
  sampleBatch(xy) { return native().stockSample(this.handle, Float64Array.from(xy)); }
  floor(x, y) { const s = this.sample(x, y); return s === -Infinity ? this.zBottom : s; }

  dispose() { native().stockDestroy(this.handle); this.handle = null; }
}
