// Evaluates Node's own Math.* on a binary input file (Float64 little-endian) so the
// oracle's restated fdlibm can be compared with the engine bit for bit.
// usage: node node_math.mjs <fn> <x.bin> <y.bin|-> <out.bin>
import fs from 'fs';
const [fn, xf, yf, of] = process.argv.slice(2);
const rd = f => { const b = fs.readFileSync(f); return new Float64Array(b.buffer, b.byteOffset, b.length / 8); };
const x = rd(xf), y = yf === '-' ? null : rd(yf), out = new Float64Array(x.length);
const F = {
  sin: i => Math.sin(x[i]), cos: i => Math.cos(x[i]), tan: i => Math.tan(x[i]),
  pow: i => Math.pow(x[i], y[i]), sq: i => x[i] ** 2, hypot: i => Math.hypot(x[i], y[i]),
  atan: i => Math.atan(x[i]), round: i => Math.round(x[i]), fmod: i => x[i] % y[i],
  atan2: i => Math.atan2(x[i], y[i]), acos: i => Math.acos(x[i]),
}[fn];
for (let i = 0; i < x.length; i++) out[i] = F(i);
fs.writeFileSync(of, Buffer.from(out.buffer));
