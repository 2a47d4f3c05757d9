// Writes tests/golden/sounding_golden.json: the reference's own Sounding.parseSharppy /
// analyze / baseState (sounding.js) evaluated on its five sample soundings (samples.js)
// and on the analytic WK sounding text, at the vertical grids the workloads use.
// Run from the repo root:  node tests/golden/make_sounding_golden.js /root/reference
'use strict';
const fs = require('fs');
const path = require('path');
const vm = require('vm');
const ref = process.argv[2] || '/root/reference';
const ctx = { window: {}, console, Math, Float32Array, Float64Array, Array, Number, isFinite, Error };
vm.createContext(ctx);
vm.runInContext(fs.readFileSync(path.join(ref, 'sounding.js'), 'utf8') + '\nthis.Sounding = Sounding;', ctx);
vm.runInContext(fs.readFileSync(path.join(ref, 'samples.js'), 'utf8'), ctx);
const wk = fs.readFileSync(path.join(__dirname, 'wk_sounding.txt'), 'utf8');
const cases = ctx.window.SAMPLE_SOUNDINGS.map(s => ({ name: s.name, text: s.text }));
cases.push({ name: 'WK analytic (th0 300, thtr 343, qv0 14 g/kg, shear 30)', text: wk });
const grids = [[40, 500], [80, 250], [100, 200]];
const out = [];
for (const c of cases) {
  const snd = ctx.Sounding.parseSharppy(c.text);
  const A = ctx.Sounding.analyze(snd);
  const rec = { name: c.name, text: c.text, nlev: snd.levels.length,
    P: { sbcape: A.P.sbcape, mlcape: A.P.mlcape, mucape: A.P.mucape, sbcin: A.P.sbcin, mlcin: A.P.mlcin },
    wind: [0, 0.5, 1, 3, 6, 10].map(z => A.windAt(z)), base: {} };
  for (const [nz, dz] of grids) {
    const b = ctx.Sounding.baseState(snd, nz, dz, [0, 0]);
    rec.base[`${nz}x${dz}`] = Array.from(b.B);
  }
  out.push(rec);
}
fs.writeFileSync(path.join(__dirname, 'sounding_golden.json'), JSON.stringify(out));
console.log('wrote', out.length, 'cases');
