// The learned bracket surrogate from JavaScript: a model file written by the Python side
// (mcsim.lib.Mlp.save: the network's weights, its input / output scaling and the held-out
// spread sigma of ln alim) loaded onto the device; predict() gives ln alim of stability
// points from their surrogate features (mc_surrogate_features: modes, tool, speed and the
// analytic zero-order limit).  No reference counterpart: it steers timeDomainLimitsGuided.
import { readFileSync } from 'fs';
import { native } from './native.js';
import { packCells } from './pack.js';

const MAGIC = 'MCSURR01';

export class Surrogate {
  // bytes: the model file's contents (ArrayBuffer or a Node Buffer)
  constructor(bytes) {
    const u8 = bytes instanceof ArrayBuffer ? new Uint8Array(bytes) : new Uint8Array(bytes.buffer, bytes.byteOffset, bytes.byteLength);
    if (String.fromCharCode(...u8.subarray(0, 8)) !== MAGIC) throw new Error('not a surrogate model file');
    const buf = u8.slice().buffer, dv = new DataView(buf);
    const nparams = dv.getInt32(8, true), nin = dv.getInt32(12, true);
    if (nparams !== native().mlpNparams() || nin !== native().sizes().MC_SURR_IN)
      throw new Error('model file for another network shape');
    this.sigma = dv.getFloat32(24, true);
    const f = new Float32Array(buf, 32, 2 * nin + nparams);
    this._h = native().mlpCreate(f.subarray(2 * nin), f.subarray(0, nin), f.subarray(nin, 2 * nin),
      dv.getFloat32(16, true), dv.getFloat32(20, true));
  }

  static load(path) { return new Surrogate(readFileSync(path)); }

  // ln alim predicted for packed mc_cell records
  predictCells(cells) {
    return Float64Array.from(native().mlpPredict(this._h, native().surrogateFeatures(cells)));
  }

  // ln alim predicted for stability points [{tool, modal, mat, S, arc, fz}] (revs 120, nonlinear)
  predict(points, opts) {
    const o = opts || {};
    const revs = o.revs === undefined ? 120 : o.revs, nonlinear = o.nonlinear === undefined ? true : o.nonlinear;
    return this.predictCells(packCells(points.map(p => ({ ...p, revs, nonlinear }))));
  }
}
