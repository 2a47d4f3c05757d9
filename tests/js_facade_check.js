// CPU checks of the Node facade (run by tests/test_js.py): the addon loads and exports its
// functions, the JS base state equals the reference sounding.js golden tables, and the
// uniform block for a given model state is printed for comparison with the Python host.
'use strict';
const fs = require('fs');
const path = require('path');
const O = require(path.join(__dirname, '..', 'outbreaklab_amd', 'js', 'outbreak.js'));
const G = JSON.parse(fs.readFileSync(path.join(__dirname, 'golden', 'sounding_golden.json'), 'utf8'));
let mismatches = 0;
for (const c of G) {
  for (const k in c.base) {
    const [nz, dz] = k.split('x').map(Number);
    const B = O.Sounding.baseState(O.Sounding.parseSharppy(c.text), nz, dz, [0, 0]).B;
    for (let i = 0; i < B.length; i++) if (B[i] !== Math.fround(c.base[k][i])) mismatches++;
  }
}
// uniform block of a 128 x 128 x 40 (2 km) Thompson model at t = 42 s, dt 6 s
const m = new O.RegionalModel();
Object.assign(m, { nx: 128, ny: 128, nz: 40, dx: 2000, dz: 500, Lx: 256000, Ly: 256000, pad: 8 });
m.params.spongeW = 20480; m.params.scheme = 'thompson'; m.params.noise = 0.1; m.params.lat = 35.9;
m.U.set('pad', 8, 144, 144, 0); m.U.set('dims', 128, 128, 40, 2000);
const wk = O.Sounding.baseState(O.Sounding.parseSharppy(fs.readFileSync(path.join(__dirname, 'golden', 'wk_sounding.txt'), 'utf8')), 40, 500, [0, 0]);
m.rhoSfc = wk.B[2]; m.time = 42;
const u = Array.from(m.writeUniforms(6));
let perr = '';
try { new O.PoissonDirect(30, 30, 10, 1000, 500); } catch (e) { perr = e.message; }
// the surrogate's training targets for two teacher rows (Nr 0 -> 3e4, 5e3 -> 1e3)
const rows = new Float32Array(48);
rows[9] = 0; rows[21] = 3e4; rows[24 + 9] = 5e3; rows[24 + 21] = -4e3;
const targets = Array.from(O.Surrogate.targets(rows));
// the radar host: seeded draws, a radar's layout and a cut's RADAR_FIELDS block on the same model
const rng = [7, 'KTLX', 123456789].map(sd => { const r = new O.Rng(sd); return [r.u(), r.u(), r.u()]; });
const geo = new O.Geo(35.9, -97.0, m.Lx, m.Ly, 300);
const ll = geo.fromModel(0.5 * m.Lx - 40e3, 0.5 * m.Ly - 30e3);
const R = O.RadarSim.plan(m, geo, { id: 'KTST', lat: ll[0], lon: ll[1], h: 350 }, { vcp: 212, seed: 7 });
R.vol = 3; R.az0 = R.rand();
// hurricane mode on the same model: the uniform block after three steps, and the vortex table
Object.assign(m, { hur: { ck: 1.3e-3, cdCap: 2.0e-3, dissHeat: true, landMoist: 0.3, sfcP: 101000, tauRad: 43200, radCap: 2 / 86400,
  blDepth: 1000, u10f: 0.8, gust: 1.5, thermals: 0.4 / 300, maritime: true, centre: [128000, 128000], thermalR: 200000, envR: 100000 },
  prep: true, steps: 3, phiScale: 1 / 0.6, base: wk });
m.params.lat = 20;
const hurU = Array.from(m.writeUniforms(6));
const vx = O.vortexTable(40, 500, 2000, 20, wk.B, { x: 128000, y: 128000, vmax: 30, rmw: 30000, r0: 120000, ztop: 16000, moistR: 60000, moistRH: 0.9 });
const vortex = { info: vx.info, tab: Array.from(vx.tab) };
const radar = { ll, xy: geo.fwd(36.4, -96.1), phase: R.phase, az0: R.az0, bytes: R.bytes, ngRef: R.ngRef, ngDop: R.ngDop,
  t0: R.cuts.map(c => c.t0), block: Array.from(O.RadarSim.block(m, geo, R, R.cuts[5], 123.5, 2)) };
console.log(JSON.stringify({ exports: Object.keys(O.addon).sort(), mismatches, uniforms: u, poissonError: perr, targets, rng, radar, hurU, vortex,
  levels: [O.levelsFor(2000), O.levelsFor(500), O.levelsFor(250)], pads: [O.gridPad(2000), O.gridPad(500), O.gridPad(250)] }));
