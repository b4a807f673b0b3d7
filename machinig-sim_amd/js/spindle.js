// Speed-dependent tool-tip modes (modesAt and its helpers, src/spindle.js:16-92), host side.
// Spindle bearings soften with speed: k_s(S) = k_s0 / (1 + c (S/Sref)^2) with c = d/(1-d)
// for a stiffness loss d at Sref; the tip sees it through its series chain, so each mode
// is scaled by f = k_tip(S)/k_tip(0) in stiffness and sqrt(f) in frequency.  Measured
// FRFs at several speeds replace the model (linear interpolation between the speeds, by
// mode order).  The arithmetic follows the reference's operation order, so the modes are
// the same doubles on the same JavaScript engine.

export const SOFT_DEFAULT = { drop: 0.25, Sref: 30000 };

const has = a => Array.isArray(a) && a.length > 0;

// k_s(S) / k_s0
export function bearingFactor(S, soft = SOFT_DEFAULT) {
  const raw = soft === null || soft === undefined || soft.drop === null || soft.drop === undefined ? 0 : soft.drop;
  const d = Math.min(0.9, Math.max(0, raw));
  if (!d || !(S > 0)) return 1;
  const ratio = S / (soft.Sref || 30000);
  return 1 / (1 + d / (1 - d) * ratio ** 2);
}

// k_tip(S) / k_tip(0) for a tool of stiffness kTool on a spindle of stiffness kSp [N/mm]
export function tipFactor(S, kTool, kSp, soft = SOFT_DEFAULT) {
  const b = bearingFactor(S, soft);
  if (b === 1 || !(kTool > 0) || !(kSp > 0)) return 1;
  const chain0 = 1 / kTool + 1 / kSp;
  return chain0 / (1 / kTool + 1 / (kSp * b));
}

// two lumped masses (spindle nose mSp on kSp, tool mTool on kTool): the tip's two modes,
// low to high, each {k [N/mm], fn [Hz], zeta}
export function twoDofModes(kTool, mTool, kSp, mSp, zTool = 0.03, zSp = 0.05) {
  const ks = kSp * 1e3, kt = kTool * 1e3;
  // det(K - w^2 M) = 0 as a quadratic in w^2
  const qa = mSp * mTool, qb = -((ks + kt) * mTool + kt * mSp), qc = ks * kt;
  const root = Math.sqrt(Math.max(0, qb * qb - 4 * qa * qc));
  const modes = [];
  for (const w2 of [(-qb - root) / (2 * qa), (-qb + root) / (2 * qa)]) {
    const tip = (ks + kt - w2 * mSp) / kt;          // tip amplitude with the spindle's at 1
    const mr = mSp + mTool * tip * tip;
    modes.push({ k: mr * w2 / (tip * tip) / 1e3, fn: Math.sqrt(w2) / (2 * Math.PI),
      zeta: mTool * tip * tip > mSp ? zTool : zSp });
  }
  return modes;
}

const scaled = (list, f) => list.map(m => ({ k: m.k * f, fn: m.fn * Math.sqrt(f), zeta: m.zeta }));

function blend(a, b, t) {
  if (a.length !== b.length) return t < 0.5 ? a : b;
  return a.map((m, i) => ({ k: m.k + (b[i].k - m.k) * t, fn: m.fn + (b[i].fn - m.fn) * t,
    zeta: m.zeta + (b[i].zeta - m.zeta) * t }));
}

// the {x, y} mode lists at spindle speed S
export function modesAt(modal, S, chain, soft = SOFT_DEFAULT) {
  if (has(modal.speedFrfs)) {
    const byS = modal.speedFrfs.slice().sort((p, q) => p.S - q.S);
    const axes = e => ({ x: e.modesX, y: has(e.modesY) ? e.modesY : e.modesX });
    const lo = byS[0], hi = byS[byS.length - 1];
    if (S <= lo.S) return axes(lo);
    if (S >= hi.S) return axes(hi);
    let j = 1;
    while (byS[j].S < S) j++;
    const a = axes(byS[j - 1]), b = axes(byS[j]), t = (S - byS[j - 1].S) / (byS[j].S - byS[j - 1].S);
    return { x: blend(a.x, b.x, t), y: blend(a.y, b.y, t) };
  }
  if (modal.auto && chain && chain.twoDof) {
    const td = chain.twoDof;
    const m = twoDofModes(chain.kTool, td.mTool, chain.kSp * bearingFactor(S, soft), td.mSp, modal.zeta, td.zSp);
    return { x: m, y: m };
  }
  const x = has(modal.modesX) ? modal.modesX : [{ k: modal.k, fn: modal.fn, zeta: modal.zeta }];
  const y = has(modal.modesY) ? modal.modesY : x;
  const f = chain ? tipFactor(S, chain.kTool, chain.kSp, soft) : 1;
  return f === 1 ? { x, y } : { x: scaled(x, f), y: scaled(y, f) };
}

// whether the dynamics change with speed (then the lobe diagram is built per speed)
export const speedDependent = (modal, soft) => has(modal.speedFrfs) ||
  (soft === null || soft === undefined || soft.drop === null || soft.drop === undefined ? 0 : soft.drop) > 0;
