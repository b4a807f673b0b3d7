// The stability-sweep worker protocol of src/stabworker.js:5-14 over the batched GPU
// search.  In: {key, points: [{S, modal}], tool, mat, arc, fz, nonlinear}.  Out: one
// {key, S, alim, open} per point, then {key, done: true}.  A newer key cancels the
// reporting of an older one.  Usable as a worker_threads worker (parentPort) or directly
// through handle(msg, post).
import { timeDomainLimitsAsync } from './stabtd.js';

let current = null;

export async function handle(m, post) {
  current = m.key;
  const pts = m.points.map(p => ({ tool: m.tool, modal: p.modal, mat: m.mat, S: p.S, arc: m.arc, fz: m.fz }));
  const res = await timeDomainLimitsAsync(pts, { nonlinear: m.nonlinear });
  if (current !== m.key) return;
  res.forEach((r, i) => post({ key: m.key, S: m.points[i].S, alim: r.alim, open: r.bracket[1] === Infinity }));
  post({ key: m.key, done: true });
}

async function attach() {
  let wt = null;
  try { wt = await import('worker_threads'); } catch (e) { return; }
  if (wt && !wt.isMainThread && wt.parentPort) {
    wt.parentPort.on('message', m => { handle(m, r => wt.parentPort.postMessage(r)); });
  }
}
attach();
