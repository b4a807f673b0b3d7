// Zero-order stability lobes (Altintas & Budak 1995) with the reference's exported
// signatures (src/sld.js:6 directionalCoeffs, :23 stabilityLobes, :64 alimAt).  Analytic
// and cheap: evaluated on the host, once per diagram, with the same arithmetic order as the
// reference so the lobes are identical.
import { modeSet, frf } from './frf.js';
import { native } from './native.js';
import { LOBE_ARGS_BYTES, writeLobeArgs } from './pack.js';

// averaged directional factors of the x/y force for an immersion [phiSt, phiEx]
export function directionalCoeffs(Kr, phiSt, phiEx) {
  const at = p => {
    const c2 = Math.cos(2 * p), s2 = Math.sin(2 * p);
    return [0.5 * (c2 - 2 * Kr * p + Kr * s2), 0.5 * (-s2 - 2 * p + Kr * c2),
      0.5 * (-s2 + 2 * p + Kr * c2), 0.5 * (-c2 - 2 * Kr * p - Kr * s2)];
  };
  const e = at(phiEx), s = at(phiSt);
  return { xx: e[0] - s[0], xy: e[1] - s[1], yx: e[2] - s[2], yy: e[3] - s[3] };
}

function eigenLimit(a0r, a0i, a1r, a1i, N, Ktc) {
  // roots of a0 L^2 + a1 L + 1 = 0; the critical one gives the smallest positive depth
  const dr = (a1r * a1r - a1i * a1i) - 4 * a0r, di = (a1r * a1i + a1i * a1r) - 4 * a0i;
  const mod = Math.hypot(dr, di);
  const sr = Math.sqrt((mod + dr) / 2), si = Math.sign(di || 1) * Math.sqrt(Math.max(0, (mod - dr) / 2));
  let best = null;
  for (const sg of [1, -1]) {
    const nr = -(a1r + sg * sr), ni = -(a1i + sg * si);
    const qr = 2 * a0r, qi = 2 * a0i, q = qr * qr + qi * qi;
    const Lr = (nr * qr + ni * qi) / q, Li = (ni * qr - nr * qi) / q;
    if (Lr >= 0) continue;
    const kap = Li / Lr;
    const alim = -2 * Math.PI * Lr * (1 + kap * kap) / (N * Ktc);
    if (alim > 0 && (!best || alim < best.alim)) best = { alim, kap };
  }
  return best;
}

// params: {N, Ktc, Krc, phiSt, phiEx, lobes = 8, rpmMin = 1000, rpmMax = 40000, ...modal}
export function stabilityLobes(params) {
  const N = params.N, Ktc = params.Ktc;
  const lobes = params.lobes === undefined ? 8 : params.lobes;
  const rpmMin = params.rpmMin === undefined ? 1000 : params.rpmMin;
  const rpmMax = params.rpmMax === undefined ? 40000 : params.rpmMax;
  const dc = directionalCoeffs(params.Krc / Ktc, params.phiSt, params.phiEx);
  const det = dc.xx * dc.yy - dc.xy * dc.yx;
  const ms = modeSet(params);
  const fns = ms.x.concat(ms.y).map(m => m.fn);
  const f0 = 0.8 * Math.min.apply(null, fns), f1 = 1.7 * Math.max.apply(null, fns);
  const NS = 400 * Math.min(4, ms.x.length + ms.y.length - 1);
  const pts = [];
  for (let s = 0; s < NS; s++) {
    const fc = f0 + (f1 - f0) * s / (NS - 1), wc = 2 * Math.PI * fc;
    const Px = frf(ms.x, fc), Py = frf(ms.y, fc);
    const pr = Px[0] * Py[0] - Px[1] * Py[1], pi = Px[0] * Py[1] + Px[1] * Py[0];
    const a0r = pr * det - pi * 0, a0i = pr * 0 + pi * det;
    const a1r = dc.xx * Px[0] + dc.yy * Py[0], a1i = dc.xx * Px[1] + dc.yy * Py[1];
    const best = eigenLimit(a0r, a0i, a1r, a1i, N, Ktc);
    if (best) pts.push({ wc, alim: best.alim, psi: Math.atan(best.kap) });
  }
  const out = [];
  for (let j = 0; j < lobes; j++) {
    const lobe = [];
    for (const p of pts) {
      const T = (Math.PI - 2 * p.psi + 2 * j * Math.PI) / p.wc, rpm = 60 / (N * T);
      if (rpm >= rpmMin * 0.8 && rpm <= rpmMax * 1.2) lobe.push({ rpm, alim: p.alim });
    }
    lobe.sort((a, b) => a.rpm - b.rpm);
    if (lobe.length) out.push(lobe);
  }
  return out;
}

// limiting depth at rpm: the lowest of the lobes covering it, linearly interpolated
export function alimAt(lobes, rpm) {
  let best = Infinity;
  for (const lobe of lobes) {
    for (let i = 1; i < lobe.length; i++) {
      const a = lobe[i - 1], b = lobe[i];
      if (rpm < a.rpm || rpm > b.rpm) continue;
      const t = (rpm - a.rpm) / (b.rpm - a.rpm || 1);
      best = Math.min(best, a.alim + (b.alim - a.alim) * t);
    }
  }
  return best;
}

// The composite curve Simulator.lobes builds for a speed-dependent tool tip (src/sim.js:270-281):
// one diagram per `node` min^-1 of speed, from the modal set at the node's centre
// (modalAt(S), the simulator's modalNow), read every `step` min^-1 within its own node.
// base: stabilityLobes' other parameters (N, Ktc, Krc, phiSt, phiEx, lobes, rpmMin, rpmMax).
// Returns [curve] with curve = [{rpm, alim}], 1e3 where no lobe covers a speed.  All nodes
// go to the library in one call (mc_composite_lobes, host threads).
export function compositeLobes(base, modalAt, node = 1000, step = 25) {
  const lo = base.rpmMin * 0.8, hi = base.rpmMax * 1.1;
  const starts = [];
  for (let S0 = lo; S0 < hi; S0 += node) starts.push(S0);
  const dv = new DataView(new ArrayBuffer(starts.length * LOBE_ARGS_BYTES));
  starts.forEach((S0, i) => writeLobeArgs(dv, i * LOBE_ARGS_BYTES,
    Object.assign({}, modalAt(S0 + node / 2), base, { rpmMin: S0, rpmMax: S0 + node })));
  const r = native().compositeLobes(dv.buffer, step, hi), n = r.length / 2;
  const curve = new Array(n);
  for (let i = 0; i < n; i++) curve[i] = { rpm: r[i], alim: r[n + i] };
  return [curve];
}
