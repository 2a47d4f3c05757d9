// CPU checks of the Node facade (run by tests/test_js.py): the addon loads and exports its functions, and for a
// given model state the facade's uniform block, surrogate targets, radar host and vortex table are printed for
// comparison with the Python host. argv[2]: a JSON file from the Python side with the golden WK base state
// (40 levels of 500 m) and a radar site's position in the Python host's projection.
'use strict';
const fs = require('fs');
const path = require('path');
const O = require(path.join(__dirname, '..', 'outbreaklab_amd', 'js', 'outbreak.js'));
const IN = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const B = Float32Array.from(IN.B);
// uniform block of a 128 x 128 x 40 (2 km) Thompson model at t = 42 s, dt 6 s
const m = new O.RegionalModel();
Object.assign(m, { nx: 128, ny: 128, nz: 40, dx: 2000, dz: 500, Lx: 256000, Ly: 256000, pad: 8 });
m.params.spongeW = 20480; m.params.scheme = 'thompson'; m.params.noise = 0.1; m.params.lat = 35.9;
m.U.set('pad', 8, 144, 144, 0); m.U.set('dims', 128, 128, 40, 2000);
m.rhoSfc = B[2]; m.time = 42;
const u = Array.from(m.writeUniforms(6));
let perr = '';
try { new O.PoissonDirect(30, 30, 10, 1000, 500); } catch (e) { perr = e.message; }
let serr = '';
try { m.setSounding('%RAW%\n%END%'); } catch (e) { serr = e.message; }
// the surrogate's training targets for two teacher rows (Nr 0 -> 3e4, 5e3 -> 1e3)
const rows = new Float32Array(48);
rows[9] = 0; rows[21] = 3e4; rows[24 + 9] = 5e3; rows[24 + 21] = -4e3;
const targets = Array.from(O.Surrogate.targets(rows));
// the radar host: seeded draws, a radar's layout and a cut's RADAR_FIELDS block on the same model, with the
// application's projection standing in as the site's model position
const rng = [7, 'KTLX', 123456789].map(sd => { const r = new O.Rng(sd); return [r.u(), r.u(), r.u()]; });
const geo = { lat0: 35.9, lon0: -97.0, groundMSL: 300, toModel: () => [IN.site.sx, IN.site.sy] };
const R = O.RadarSim.plan(m, geo, { id: 'KTST', lat: IN.site.lat, lon: IN.site.lon, h: 350 }, { vcp: 212, seed: 7 });
R.vol = 3; R.az0 = R.rand();
// hurricane mode on the same model: the uniform block after three steps, and the vortex table
Object.assign(m, { hur: { ck: 1.3e-3, cdCap: 2.0e-3, dissHeat: true, landMoist: 0.3, sfcP: 101000, tauRad: 43200, radCap: 2 / 86400,
  blDepth: 1000, u10f: 0.8, gust: 1.5, thermals: 0.4 / 300, maritime: true, centre: [128000, 128000], thermalR: 200000, envR: 100000 },
  prep: true, steps: 3, phiScale: 1 / 0.6, base: { B } });
m.params.lat = 20;
const hurU = Array.from(m.writeUniforms(6));
const vx = O.vortexTable(40, 500, 2000, 20, B, { x: 128000, y: 128000, vmax: 30, rmw: 30000, r0: 120000, ztop: 16000, moistR: 60000, moistRH: 0.9 });
const vortex = { info: vx.info, tab: Array.from(vx.tab) };
const radar = { phase: R.phase, az0: R.az0, bytes: R.bytes, ngRef: R.ngRef, ngDop: R.ngDop, t0: R.cuts.map(c => c.t0),
  dur: R.cuts.map(c => c.dur), offs: R.cuts.map(c => c.mom.map(x => x.off)), block: Array.from(O.RadarSim.block(m, geo, R, R.cuts[5], 123.5, 2)) };
console.log(JSON.stringify({ exports: Object.keys(O.addon).sort(), uniforms: u, poissonError: perr, soundingError: serr, targets, rng, radar, hurU,
  vortex, gridBytes: O.gridBytes({ dx: 500, n: [1024, 1024, 80] }),
  levels: [2000, 500, 250, 750, 120, 5000].map(O.levelsFor), pads: [2000, 500, 250, 100].map(O.gridPad),
  custom: [O.customGrid({ dx: 500, w: 512, h: 512 }), O.customGrid({ dx: 2000, w: 5000, h: 10 })] }));
