// Time-domain stability limit with the reference's exported signatures
// (src/stabtd.js:14 chatterAt, :24 timeDomainLimit), computed on the MI355X by the
// persistent chatter-search kernel: same trials, same verdicts (threshold 0.45), same
// bracket as the JavaScript engine, bit for bit.
import { native } from './native.js';
import { packCells, readLimits, readTrials } from './pack.js';

export const CHATTER_IDX = 0.45;

function cellOf(tool, modal, mat, S, arc, fz, revs, nonlinear) {
  return { tool, modal, mat, S, arc, fz, revs, nonlinear };
}

// tool: {N, R, helixDeg, type, fluteLen}; modal: modes at S; mat; arc: [phiSt, phiEx];
// fz [mm]; ap [mm].  Returns true when the cut chatters.
export function chatterAt(tool, modal, mat, S, arc, fz, ap, opts) {
  const o = opts || {};
  const revs = o.revs === undefined ? 120 : o.revs, nonlinear = o.nonlinear === undefined ? true : o.nonlinear;
  const cells = packCells([cellOf(tool, modal, mat, S, arc, fz, revs, nonlinear)]);
  return readTrials(native().chatterBatch(cells, new Float64Array([ap])), 1)[0].chatter;
}

// Returns { alim, bracket: [stable, chatter] } (bracket [aMax, Infinity] when stable at aMax).
export function timeDomainLimit(tool, modal, mat, S, arc, fz, opts) {
  return timeDomainLimits([{ tool, modal, mat, S, arc, fz }], opts)[0];
}

function optsOf(opts) {
  const o = opts || {};
  return { aMin: o.aMin === undefined ? 0.05 : o.aMin, aMax: o.aMax === undefined ? 20 : o.aMax,
    iters: o.iters === undefined ? 8 : o.iters, revs: o.revs === undefined ? 120 : o.revs,
    nonlinear: o.nonlinear === undefined ? true : o.nonlinear };
}

const asResult = r => ({ alim: r.alim, bracket: [r.lo, r.open ? Infinity : r.hi], trials: r.trials, steps: r.steps });

// Batch form: points = [{tool, modal, mat, S, arc, fz}], one launch for all of them.
export function timeDomainLimits(points, opts) {
  const o = optsOf(opts);
  const cells = packCells(points.map(p => cellOf(p.tool, p.modal, p.mat, p.S, p.arc, p.fz, o.revs, o.nonlinear)));
  return readLimits(native().bisectBatch(cells, o.aMin, o.aMax, o.iters), points.length).map(asResult);
}

// Asynchronous batch form (a worker thread runs the launch; the event loop stays free).
export function timeDomainLimitsAsync(points, opts) {
  const o = optsOf(opts);
  const cells = packCells(points.map(p => cellOf(p.tool, p.modal, p.mat, p.S, p.arc, p.fz, o.revs, o.nonlinear)));
  return native().bisectBatchAsync(cells, o.aMin, o.aMax, o.iters).then(buf => readLimits(buf, points.length).map(asResult));
}

// Bracketed search from a prior (e.g. the surrogate): lo/hi per point, stop at width tol.
export function timeDomainLimitsBracketed(points, lo, hi, tol, opts) {
  const o = optsOf(opts), n = points.length;
  const cells = packCells(points.map(p => cellOf(p.tool, p.modal, p.mat, p.S, p.arc, p.fz, o.revs, o.nonlinear)));
  const buf = native().bracketBatch(cells, o.aMin, o.aMax, o.iters, Float64Array.from(lo), Float64Array.from(hi), tol);
  const lim = readLimits(buf, n), fb = new Int32Array(buf, n * 56, n);
  return lim.map((r, i) => Object.assign(asResult(r), { fellBack: fb[i] !== 0 }));
}

// timeDomainLimits steered by the surrogate: the same results as timeDomainLimits (same
// trials on the search path, same verdicts, same brackets), with the speculative search
// ordered by the prior ln alim ~ N(prediction, sigma^2) so that fewer speculative trials run
// when the prediction is good.  surrogate: a Surrogate (surrogate.js); opts.sigma overrides
// the spread stored with the model.
export function timeDomainLimitsGuided(points, surrogate, opts) {
  const o = optsOf(opts), n = points.length;
  const cells = packCells(points.map(p => cellOf(p.tool, p.modal, p.mat, p.S, p.arc, p.fz, o.revs, o.nonlinear)));
  const mu = surrogate.predictCells(cells);
  const s = opts && opts.sigma !== undefined ? opts.sigma : surrogate.sigma;
  const sig = new Float64Array(n).fill(s > 0 ? s : 0.3);
  return readLimits(native().bisectPriorBatch(cells, o.aMin, o.aMax, o.iters, mu, sig), n).map(asResult);
}
