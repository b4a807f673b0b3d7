// Node 12 ES-module loader hook used ONLY to generate golden fixtures in the build
// container.  The reference's hot-path modules use `a ?? b` and `a?.b`, which V8 7.8
// (Node 12) cannot parse.  This hook rewrites exactly those two operators, in memory,
// into calls of two helpers with the same semantics (`__nc(a, b)` = a ?? b for
// side-effect-free b, `__oc(a, 'k')` = a?.k, `__oc(a, e)` = a?.[e]) and leaves every other
// token untouched.
// Nothing is written to disk.
const REF = process.env.MC_REFERENCE_SRC || '/root/reference/src/';

function scanLeftMember(s, i) {            // start index of the member chain ending at i
  let depth = 0, j = i;
  for (; j >= 0; j--) {
    const c = s[j];
    if (c === ')' || c === ']') { depth++; continue; }
    if (c === '(' || c === '[') { if (depth === 0) break; depth--; continue; }
    if (depth > 0) continue;
    if (/[\w$.]/.test(c)) continue;
    break;
  }
  return j + 1;
}

function rewriteOptional(s) {
  let k;
  while ((k = s.indexOf('?.')) >= 0) {
    if (/\d/.test(s[k + 2] || '')) break;                 // a ternary `?.5`: not optional chaining
    const start = scanLeftMember(s, k - 1);
    if (s[k + 2] === '[') {                               // a?.[expr]
      let depth = 0, e = k + 2;
      for (; e < s.length; e++) { if (s[e] === '[') depth++; else if (s[e] === ']' && --depth === 0) break; }
      const left = s.slice(start, k), key = s.slice(k + 3, e);
      s = s.slice(0, start) + `__oc(${left}, ${key})` + s.slice(e + 1);
      continue;
    }
    const m = s.slice(k + 2).match(/^[A-Za-z_$][\w$]*/);
    if (!m) throw new Error('unsupported ?. form near: ' + s.slice(start, k + 12));
    const left = s.slice(start, k);
    s = s.slice(0, start) + `__oc(${left}, '${m[0]}')` + s.slice(k + 2 + m[0].length);
  }
  return s;
}

function scanLeftOperand(s, i) {
  let depth = 0, j = i;
  for (; j >= 0; j--) {
    const c = s[j];
    if (c === ')' || c === ']' || c === '}') { depth++; continue; }
    if (c === '(' || c === '[' || c === '{') { if (depth === 0) break; depth--; continue; }
    if (depth > 0) continue;
    if (c === ',' || c === ';' || c === ':' || c === '?' || c === '\n') break;
    if (c === '=' && !/[=!<>]/.test(s[j - 1] || '') && s[j + 1] !== '=' && s[j + 1] !== '>') break;
  }
  return j + 1;
}

function scanRightOperand(s, i) {
  let depth = 0, j = i;
  for (; j < s.length; j++) {
    const c = s[j];
    if (c === "'" || c === '"') { j = s.indexOf(c, j + 1); if (j < 0) j = s.length; continue; }   // a string literal
    if (c === '(' || c === '[' || c === '{') { depth++; continue; }
    if (c === ')' || c === ']' || c === '}') { if (depth === 0) break; depth--; continue; }
    if (depth > 0) continue;
    if (c === ',' || c === ';' || c === ':' || c === '\n') break;
    if (c === '?' ) break;
  }
  return j;
}

function rewriteNullish(s) {
  let k;
  while ((k = s.indexOf('??')) >= 0) {
    const ls = scanLeftOperand(s, k - 1), re = scanRightOperand(s, k + 2);
    const left = s.slice(ls, k).trim(), right = s.slice(k + 2, re).trim();
    const lead = s.slice(ls, k).match(/^\s*/)[0];
    s = s.slice(0, ls) + lead + `__nc(${left}, ${right})` + s.slice(re);
  }
  return s;
}

export function downlevel(src) {
  const lines = String(src).split('\n');
  const out = lines.map(l => {
    const code = l.replace(/\/\/.*$/, m => (/['"`]/.test(m) ? m : ''));
    if (!code.includes('??') && !code.includes('?.')) return l;
    return rewriteNullish(rewriteOptional(code));
  });
  out.push('function __oc(o, k) { return o === null || o === undefined ? undefined : o[k]; }');
  out.push('function __nc(a, b) { return a === null || a === undefined ? b : a; }');
  return out.join('\n');
}

export async function transformSource(source, context, defaultTransformSource) {
  if (context.url.startsWith('file://' + REF)) return { source: downlevel(source) };
  return defaultTransformSource(source, context, defaultTransformSource);
}
