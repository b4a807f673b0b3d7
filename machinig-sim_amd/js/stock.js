// Z-heightfield stock with the reference's interface (src/stock.js:8-165), kept resident
// in MI355X HBM.  stamp() / stampBatch() remove material on the device with the
// reference's sequential semantics (H, err, wallD, lineCut, lineWall bit-identical); the
// public typed arrays are host mirrors refreshed from the device on first access after a
// change.  A caller that writes H directly calls uploadH() afterwards.
import { native } from './native.js';
import { packStamps } from './pack.js';

export class Stock {
  constructor(w, l, h, cell) {
    const c = cell === undefined ? 0.3 : cell;
    this.x0 = 0; this.y0 = 0; this.w = w; this.l = l;
    this.zTop = 0; this.zBottom = -h;
    this.handle = native().stockCreate(w, l, h, c);
    const d = native().stockDims(this.handle);
    this.nx = d.nx; this.ny = d.ny; this.hx = d.hx; this.hy = d.hy;
    this.curLine = 0;
    this.removed = 0;
    this._wallBand = 0;
    this._mirror = null;
    this.clearDirty();
  }

  get wallBand() { return this._wallBand; }
  set wallBand(v) { this._wallBand = v || 0; native().stockSetWallBand(this.handle, this._wallBand); }

  _sync() {
    if (this._mirror) return this._mirror;
    const n = this.nx * this.ny;
    const m = { H: new Float32Array(n), err: new Float32Array(n), wallD: new Float32Array(n),
      lineCut: new Int32Array(n), lineWall: new Int32Array(n) };
    native().stockRead(this.handle, m.H, m.err, m.wallD, m.lineCut, m.lineWall);
    this._mirror = m;
    return m;
  }
  get H() { return this._sync().H; }
  get err() { return this._sync().err; }
  get wallD() { return this._sync().wallD; }
  get lineCut() { return this._sync().lineCut; }
  get lineWall() { return this._sync().lineWall; }

  // push a host-edited H to the device
  uploadH(H) { native().stockWriteH(this.handle, H || this._sync().H); this._mirror = null; }

  clearDirty() { this.dirty = { i0: Infinity, j0: Infinity, i1: -1, j1: -1 }; }
  inside(x, y) { return x >= 0 && y >= 0 && x <= this.w && y <= this.l; }

  // remove material under the tool ('flat' | 'ball'); returns the removed volume [mm^3]
  stamp(cx, cy, tipZ, R, type, errVal) {
    return this.stampBatch([{ cx, cy, tipZ, R, type, errVal, line: this.curLine }])[0];
  }

  // many stamps in one launch, applied in order: [{cx, cy, tipZ, R, type, errVal, line}]
  stampBatch(list) {
    if (!list.length) return new Float64Array(0);
    const vol = native().stampBatch(this.handle, packStamps(list));
    const Rmax = Math.max(2 * Math.max(this.hx, this.hy), this._wallBand);
    for (let i = 0; i < list.length; i++) {
      const s = list[i];
      this.removed += vol[i];
      if (s.tipZ >= this.zTop) continue;
      const Rw = s.R + Rmax, d = this.dirty;
      d.i0 = Math.min(d.i0, Math.max(0, Math.floor((s.cx - Rw) / this.hx)));
      d.j0 = Math.min(d.j0, Math.max(0, Math.floor((s.cy - Rw) / this.hy)));
      d.i1 = Math.max(d.i1, Math.min(this.nx - 1, Math.ceil((s.cx + Rw) / this.hx)));
      d.j1 = Math.max(d.j1, Math.min(this.ny - 1, Math.ceil((s.cy + Rw) / this.hy)));
    }
    this._mirror = null;
    return vol;
  }

  // bilinear height; outside the stock -Infinity
  sample(x, y) { return this.sampleBatch(new Float64Array([x, y]))[0]; }
  sampleBatch(xy) { return native().stockSample(this.handle, Float64Array.from(xy)); }
  floor(x, y) { const s = this.sample(x, y); return s === -Infinity ? this.zBottom : s; }

  dispose() { native().stockDestroy(this.handle); this.handle = null; }
}
