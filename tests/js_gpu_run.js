// GPU run of the Node facade (run by tests/test_gpu_js.py): the 128 x 128 x 40, 2 km Thompson
// workload driven through RegionalModel / the N-API addon for 20 steps; prints the state.
// argv[2]: the output folder; argv[3]: a JSON file from the Python side with the WK base state (40 levels of
// 500 m) that the application's Sounding module would build, and the radar site's position in the application's
// projection (the facade takes both from the application, as model.js and radar.js do).
'use strict';
const fs = require('fs');
const path = require('path');
const O = require(path.join(__dirname, '..', 'outbreaklab_amd', 'js', 'outbreak.js'));
const IN = JSON.parse(fs.readFileSync(process.argv[3], 'utf8'));
O.useSounding({ parseSharppy: text => ({ text }), analyze: () => ({ stub: true }), baseState: () => ({ B: Float32Array.from(IN.B) }) });
(async () => {
  const m = new O.RegionalModel();
  await m.allocate('custom', { dx: 2000, n: [128, 128, 40] });
  m.params.scheme = 'thompson'; m.params.noise = 0.1; m.params.lat = 35.9;
  m.setSounding(fs.readFileSync(path.join(__dirname, 'golden', 'wk_sounding.txt'), 'utf8'));
  m.setBoundaries([]); m.setCI([]); m.setSurface(null, null);
  m.reset();
  m.addBubble(0.5 * m.Lx, 0.5 * m.Ly, 1500, 10000, 1500, 2);
  for (let s = 0; s < 20; s++) {
    const enc = O.GPU.device.createCommandEncoder();
    m.encStep(enc, 6.0);
    O.GPU.device.queue.submit([enc.finish()]);
  }
  await m.sync();
  const out = path.join(process.argv[2] || '.', 'js_state.bin');
  const V = m.Vcur.download(), Q1 = m.Q1[0].download(), Q2 = m.Q2[0].download();
  fs.writeFileSync(out, Buffer.concat([Buffer.from(V.buffer), Buffer.from(Q1.buffer), Buffer.from(Q2.buffer)]));
  // requestMaps of the final state (model.js:1478-1494)
  const maps = await m.requestMaps(null)();
  const mout = path.join(process.argv[2] || '.', 'js_maps.bin');
  fs.writeFileSync(mout, Buffer.concat([0, 1, 2, 3].map(q => Buffer.from(maps['m' + q].slice().buffer))));
  // one seeded VCP 35 volume of a radar 40 km west / 30 km south of the centre, one sub-grid tornado
  const geo = { lat0: 35.9, lon0: -97.0, groundMSL: 0, toModel: () => [IN.site.sx, IN.site.sy] };
  const sim = new O.RadarSim(m, geo);
  sim.setTornadoes([{ x: 0.5 * m.Lx, y: 0.5 * m.Ly, R: 150, v: 60, debrisTop: 1200, debrisDbz: 55, debris: 1 }]);
  const R = sim.addRadar({ id: 'KTST', lat: IN.site.lat, lon: IN.site.lon, h: 0 }, { vcp: 35, seed: 11 });
  const done = sim.encDue(null, R.phase + R.vcp.dur - 1);
  const vol = await sim.collect(done[0], 'js', 0);
  const rout = path.join(process.argv[2] || '.', 'js_radar.bin');
  fs.writeFileSync(rout, Buffer.from(vol.bytes.buffer));
  sim.destroy();
  // hurricane mode: the 128 x 128 x 40 hurricane workload (configs.setup_hurricane_128) on the golden WK
  // sounding, 9 steps (the 2-dx filter at steps 3 and 7, PREP from step 1)
  const h = new O.RegionalModel();
  await h.allocate('custom', { dx: 2000, n: [128, 128, 40] });
  Object.assign(h.params, { qc0: 1e-3, dampZ: 16500, dampTau: 300, mcFloor: 1.0, evap: 0.10 });
  h.params.scheme = 'lin'; h.params.noise = 0.1; h.params.lat = 20;
  h.setSounding(fs.readFileSync(path.join(__dirname, 'golden', 'wk_sounding.txt'), 'utf8'));
  h.setBoundaries([]);
  const c = [0.5 * h.Lx, 0.5 * h.Ly];
  h.setCI([{ x: c[0] + 40000, y: c[1], zc: 2000, rh: 10000, rz: 1500, wt: 8, t0: 0, t1: 1800, ramp: 60 }]);
  const surf = new Float32Array(h.nx * h.ny * 4);
  for (let j = 0; j < h.ny; j++) for (let i = 0; i < h.nx; i++) {
    surf[(j * h.nx + i) * 4 + 1] = 302;
    if ((i + 0.5) * h.dx > 200000) surf[(j * h.nx + i) * 4] = 1;
  }
  h.setSurface(surf, { ck: 1.3e-3, cdCap: 2.0e-3, dissHeat: true, thermalFade: true, landMoist: 0.3, tauRad: 43200,
    radCap: 2 / 86400, blDepth: 1000, u10f: 0.8, gust: 1.5, thermals: 0.4 / 300, maritime: true, thermalR: 200000,
    sfcP: 101000, centre: c, envR: 100000 });
  h.setVortex({ x: c[0], y: c[1], vmax: 30, rmw: 30000, r0: 120000, ztop: 16000, moistR: 60000, moistRH: 0.9 });
  h.reset();
  for (let s = 0; s < 9; s++) h.encStep(null, 6.0);
  await h.sync();
  const hout = path.join(process.argv[2] || '.', 'js_hur.bin');
  fs.writeFileSync(hout, Buffer.concat([h.Vcur.download(), h.Q1[0].download(), h.Q2[0].download()].map(a => Buffer.from(a.buffer))));
  const hinfo = { hout, hparity: h.parity, hsteps: h.steps, hphiScale: h.phiScale };
  h.destroy();
  console.log(JSON.stringify({ out, mout, rout, hinfo, header: vol.header, volumes: done.length, mapsT: maps.t, parity: m.parity,
    time: m.time, nV: V.length, nQ: Q1.length }));
})().catch(e => { console.error(e.stack || e); process.exit(1); });
