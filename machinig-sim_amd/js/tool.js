// End mill as a uniform carbide cantilever in series with the spindle/holder spring
// (toolModel, src/tool.js:16-32), host side: tip stiffness and first bending frequency.
// Stepped beams (holders, necks, imported meshes: src/assembly.js) are not part of this
// module and raise an Error.  The arithmetic follows the reference's operation order.

const E_CARBIDE = 600000;    // MPa
const RHO_CARBIDE = 14500;   // kg/m^3

export function isAssembly(t) {
  return !!(t.holderSegs || t.stlTool || t.stlHolder || t.neckD || (t.shankD && Math.abs(t.shankD - t.D) > 1e-6));
}

// {De, I, kTool, k, fn, mEff}: De [mm] equivalent fluted diameter, I [mm^4], k [N/mm], fn [Hz]
export function toolModel(t) {
  if (isAssembly(t)) throw new Error('toolModel: stepped tool assemblies are not supported by this module');
  const De = 0.8 * t.D, L = t.stickout;
  const I = Math.PI * De ** 4 / 64;
  const kTool = 3 * E_CARBIDE * I / L ** 3;
  const k = 1 / (1 / kTool + 1 / t.kSpindle);
  // first cantilever mode of the tool alone, SI units: (1.875)^2 sqrt(EI / rhoA) / L^2
  const EI = E_CARBIDE * 1e6 * I * 1e-12;
  const rhoA = RHO_CARBIDE * Math.PI * (De * 1e-3) ** 2 / 4;
  const w = 1.875 ** 2 * Math.sqrt(EI / rhoA) / (L * 1e-3) ** 2;
  const mHolder = t.mHolder === null || t.mHolder === undefined ? 0.05 : t.mHolder;
  const mEff = kTool * 1e3 / w ** 2 + mHolder;
  return { De, I, kTool, k, fn: Math.sqrt(k * 1e3 / mEff) / (2 * Math.PI), mEff };
}
