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
  console.log(JSON.stringify({ out, mout, mapsT: maps.t, parity: m.parity, time: m.time, nV: V.length, nQ: Q1.length }));
})().catch(e => { console.error(e.stack || e); process.exit(1); });
