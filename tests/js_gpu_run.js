// GPU run of the Node facade (run by tests/test_gpu_js.py): the 128 x 128 x 40, 2 km Thompson
// workload driven through RegionalModel / the N-API addon for 20 steps; prints the state.
'use strict';
const fs = require('fs');
const path = require('path');
const O = require(path.join(__dirname, '..', 'outbreaklab_amd', 'js', 'outbreak.js'));
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
  const geo = new O.Geo(35.9, -97.0, m.Lx, m.Ly, 0);
  const sim = new O.RadarSim(m, geo);
  sim.setTornadoes([{ x: 0.5 * m.Lx, y: 0.5 * m.Ly, R: 150, v: 60, debrisTop: 1200, debrisDbz: 55, debris: 1 }]);
  const ll = geo.fromModel(0.5 * m.Lx - 40e3, 0.5 * m.Ly - 30e3);
  const R = sim.addRadar({ id: 'KTST', lat: ll[0], lon: ll[1], h: 0 }, { vcp: 35, seed: 11 });
  const done = sim.encDue(null, R.phase + R.vcp.dur - 1);
  const vol = await sim.collect(done[0], 'js', 0);
  const rout = path.join(process.argv[2] || '.', 'js_radar.bin');
  fs.writeFileSync(rout, Buffer.from(vol.bytes.buffer));
  sim.destroy();
  console.log(JSON.stringify({ out, mout, rout, header: vol.header, volumes: done.length, mapsT: maps.t, parity: m.parity,
    time: m.time, nV: V.length, nQ: Q1.length }));
})().catch(e => { console.error(e.stack || e); process.exit(1); });
