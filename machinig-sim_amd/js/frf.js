// Mode sets and receptance of the tool tip, with the reference's conventions
// (src/frf.js:6-21): a mode is {k [N/mm], fn [Hz], zeta}; modal = {k, fn, zeta} for one
// symmetric mode, or modesX / modesY lists (y defaults to x); FRF values in mm/N.

const hasModes = a => Array.isArray(a) && a.length > 0;

export function modeSet(modal) {
  const single = [{ k: modal.k, fn: modal.fn, zeta: modal.zeta }];
  const x = hasModes(modal.modesX) ? modal.modesX : single;
  const y = hasModes(modal.modesY) ? modal.modesY : x;
  return { x, y };
}

// G(f) = sum over modes of 1 / (k (1 - r^2 + 2 i zeta r)), r = f / fn
export function frf(modes, f) {
  let re = 0, im = 0;
  for (let j = 0; j < modes.length; j++) {
    const m = modes[j];
    const r = f / m.fn, a = 1 - r * r, b = 2 * m.zeta * r, d = m.k * (a * a + b * b);
    re += a / d;
    im -= b / d;
  }
  return [re, im];
}
