'use strict';
// ─────────────────────────────────────────────────────────────────────────────
//  RegionalModel / PoissonDirect for Node on MI355X.
//
//  The same class and method names, argument shapes, parity convention and
//  error texts as the reference's model.js:1156-1496 and solver.js:105-165; the
//  GPU work goes through the N-API addon (obl_napi.node) to liboutbreak_hip.so
//  (gfx950 HIP kernels). No WGSL is dispatched and there is no WebGPU fallback.
//
//  `enc` / `pass` arguments are accepted and ignored: work is enqueued on the
//  model's HIP stream in call order, as on the reference's single queue.
//  A minimal `GPU` object gives the driver loop its queue calls
//  (createCommandEncoder, queue.submit, queue.onSubmittedWorkDone).
//
//  Written in ES2019 (no ?? or ?.), so Node 12 runs it.
// ─────────────────────────────────────────────────────────────────────────────
const path = require('path');
const zlib = require('zlib');
const addon = require('./obl_napi.node');
addon.load(process.env.OBL_LIB || path.join(__dirname, '..', 'liboutbreak_hip.so'));

const DEG = Math.PI / 180;
const clamp = (v, a, b) => Math.min(Math.max(v, a), b);
const dflt = (v, d) => (v === undefined || v === null ? d : v);
const FIELD = { V: 0, Q1: 1, Q2: 2, Vt: 3, Q1t: 4, Q2t: 5, aux: 6, phT: 7, lsfA: 8, lsfB: 9, ciMap: 10, phi: 11, rhs: 12, bs: 13, sfc: 14, Vp: 15 };
const GEMM = { fp32: 0, bf16x3: 1, bf16x6: 2 };
const THOMAS = { fp32: 0, fp64: 1, mixed: 2 };
const MAX_BOUNDS = 8, MAX_BPTS = 32, MAX_CI = 256;
const CN_OLD = 0.4;  // hurricane mode: the share of the last step's pressure force applied before advection

// ── the uniform block: MODEL_FIELDS (model.js:23-37), UBlock (util.js:83-98) ──
const MODEL_FIELDS = ['dims', 'dzv', 'step', 'sponge', 'turb', 'micro', 'misc', 'cor', 'ci', 'pad', 'micro2',
  'hur', 'hur2', 'hur3', 'hur4', 'hur5', 'vtx', 'vtx2', 'cn'];
class UBlock {
  constructor(fields) {
    this.off = {};
    fields.forEach((n, i) => { this.off[n] = 4 * i; });
    this.f = new Float32Array(4 * fields.length);
    this.size = this.f.byteLength;
  }
  set(name, a, b, c, d) {
    const o = this.off[name];
    if (o === undefined) throw new Error('UBlock: no field ' + name);
    this.f[o] = dflt(a, 0); this.f[o + 1] = dflt(b, 0); this.f[o + 2] = dflt(c, 0); this.f[o + 3] = dflt(d, 0);
    return this;
  }
}

// ── grid helpers (model.js:1104-1154): level counts, custom grids, the solve's padding, device bytes ──
// (grid spacing m, levels over 20 km), model.js:1104
const LEVELS_BY_DX = [[100, 160], [250, 100], [500, 80], [1000, 64], [1500, 60], [2000, 56], [3000, 52], [4000, 48]];
const LOG_DX = LEVELS_BY_DX.map(r => Math.log(r[0]));
const nearestMultiple = (x, m) => Math.round(x / m) * m;
// levelsFor: linear in log(dx) between the table's rows, to the nearest multiple of 4; the end rows beyond it
function levelsFor(dx) {
  const l = Math.log(clamp(dx, LEVELS_BY_DX[0][0], LEVELS_BY_DX[LEVELS_BY_DX.length - 1][0]));
  const i = LOG_DX.findIndex((v, j) => j > 0 && l <= v);
  if (i < 0) return LEVELS_BY_DX[LEVELS_BY_DX.length - 1][1];
  const t = (l - LOG_DX[i - 1]) / (LOG_DX[i] - LOG_DX[i - 1]), n0 = LEVELS_BY_DX[i - 1][1], n1 = LEVELS_BY_DX[i][1];
  return nearestMultiple(n0 + (n1 - n0) * t, 4);
}
// customGrid: spacing 100-4000 m (default 500) and a domain in km (default 256) -> cells per side, the nearest
// multiple of 8 within 64..2040
function customGrid(c) {
  const dx = clamp(+c.dx || 500, 100, 4000);
  const [nx, ny] = [c.w, c.h].map(km => clamp(nearestMultiple((+km || 256) * 1000 / dx, 8), 64, 2040));
  const nz = levelsFor(dx);
  return { key: `custom:${dx}:${nx}x${ny}x${nz}`, dx, n: [nx, ny, nz] };
}
// gridPad: about 10 km of padding cells per side, the nearest multiple of 8 within 8..64
function gridPad(dx) { return clamp(nearestMultiple(10000 / dx, 8), 8, 64); }
// the device bytes the library allocates for a grid (its own layout; the counterpart of gridBytes)
function gridBytes(g) { return addon.gridBytes(g.n[0], g.n[1], g.n[2], g.dx, 20000 / g.n[2], gridPad(g.dx)); }
const GRID_PRESETS = {
  quick: { dx: 2000, n: [320, 320, 56] }, standard: { dx: 1500, n: [448, 448, 60] }, large: { dx: 1500, n: [640, 640, 60] },
  fine: { dx: 1000, n: [640, 640, 64] }, m500: { dx: 500, n: [512, 512, 80] }, m250: { dx: 250, n: [512, 512, 100] },
  m100: { dx: 100, n: [512, 512, 160] }, h2: { dx: 2000, n: [640, 640, 56] }, h3: { dx: 3000, n: [640, 640, 52] },
  h15: { dx: 1500, n: [960, 960, 60] },
};

// ── soundings: the caller's Sounding module ──
// The base state comes from the application's own sounding code (sounding.js's Sounding: parseSharppy, analyze,
// baseState), as in the reference, where model.js calls it from setSounding; this facade does not carry a copy.
// useSounding(S) hands it over once; setBaseState takes a prepared table instead.
let SOUNDING = null;
function useSounding(S) {
  if (!S || typeof S.parseSharppy !== 'function' || typeof S.baseState !== 'function') throw new Error('useSounding: a Sounding module with parseSharppy and baseState expected');
  SOUNDING = S;
}

// ── a minimal stand-in for the driver's GPU object (util.js:101-191): queue calls only ──
function makeGPU() {
  const models = new Set();
  return {
    errors: [],
    device: {
      createCommandEncoder() { return { beginComputePass() { return { end() {} }; }, finish() { return {}; } }; },
      queue: {
        submit() {},
        onSubmittedWorkDone() { return Promise.all(Array.from(models).map(m => m.h ? addon.sync(m.h) : null)); },
      },
    },
    track(m) { models.add(m); },
    untrack(m) { models.delete(m); },
  };
}
const GPU = makeGPU();

class PoissonDirect {
  // solver.js:106-137; nx, ny must be multiples of 8 on this GPU path
  constructor(nx, ny, nz, dx, dz) {
    Object.assign(this, { nx, ny, nz, dx, dz });
    this.h = addon.poissonCreate(nx, ny, nz, dx, dz);
    this.setDensity(new Array(nz).fill(1), new Array(nz + 1).fill(1));
  }
  setDensity(rhoC, rhoW) { addon.poissonSetDensity(this.h, Float64Array.from(rhoC), Float64Array.from(rhoW)); }
  // host arrays (x fastest); precision: {gemm: 'bf16x6'|'fp32'|'bf16x3', thomas: 'fp32'|'fp64'|'mixed', cond}
  solve(rhs, opt) {
    const o = opt || {};
    const phi = new Float32Array(rhs.length);
    addon.poissonSolve(this.h, rhs instanceof Float32Array ? rhs : Float32Array.from(rhs), phi,
      GEMM[o.gemm || 'bf16x6'], THOMAS[o.thomas || 'mixed'], dflt(o.cond, 64));
    return phi;
  }
  destroy() { this.h = null; }
}

// a device field: download()/upload() copy it as a typed array in the reference's texel layout
class Field {
  constructor(m, name, copy, count, Ctor) { Object.assign(this, { m, name, copy, count, Ctor: Ctor || Float32Array }); }
  download() { const a = new this.Ctor(this.count); addon.download(this.m.h, FIELD[this.name], this.copy, a); return a; }
  upload(a) { addon.upload(this.m.h, FIELD[this.name], this.copy, a); }
  get ptr() { return addon.fieldPtr(this.m.h, FIELD[this.name], this.copy); }
}

// setVortex's table (model.js:1318-1365), built by the library's host routine obl_vortex_table (the one the Python
// host calls): tangential wind v(r, z) = V(r) F(z), phi in gradient-wind balance, theta' = theta0 (d phi / dz) / g
function vortexTable(nz, dz, dx, lat, B, v) {
  const r = addon.vortexTable(nz, dz, dx, lat, Float32Array.from(B), Float64Array.from([v.vmax, v.rmw, v.r0, v.ztop]), v.shape === 'mid' ? 1 : 0);
  const [dr, nr, pc, minStab, warm] = r.info;
  return { tab: r.tab, info: { x: v.x, y: v.y, dr, nr, moistR: v.moistR || 1, moistRH: v.moistR ? v.moistRH : 0, pc, minStab, warm } };
}

class RegionalModel {
  constructor() {
    this.U = new UBlock(MODEL_FIELDS);
    // model.js:1176-1180, with the outbreak overrides of app.js:1613
    this.params = {
      cs: 0.18, nu0: 1.0, drag: 1, z0: 0.1, spongeW: 36000, spongeTau: 100, dampZ: 15000, dampTau: 150,
      auto: 1e-3, qc0: 2e-3, freeze: 1.0, noise: 0.25, mc: 1.0, vcap: 110, evap: 0.22, sortQ: 5e-4, ciRate: 0.5, nudgeTau: 5400, lat: 35,
      moistMix: 0, mcFloor: 0.25,
    };
    this.parity = 0; this.time = 0; this.steps = 0;
    this.bounds = []; this.ciEvents = []; this.ciActive = 0;
    this.mapBusy = [false, false];
    this.noiseSeed = 0;
    this.h = null;
  }
  // model.js:1186-1201: throws with a user-facing message when the grid does not fit
  async allocate(key, spec) {
    this.destroy();
    const pr = spec || GRID_PRESETS[key], [nx, ny, nz] = pr.n;
    Object.assign(this, { key, nx, ny, nz, dx: pr.dx, dz: 20000 / nz });
    this.Lx = nx * this.dx; this.Ly = ny * this.dx; this.Lz = nz * this.dz;
    this.params.spongeW = clamp(0.08 * Math.min(this.Lx, this.Ly), 5000, 40000);
    this.pad = gridPad(this.dx);
    const nxp = nx + 2 * this.pad, nyp = ny + 2 * this.pad;
    try { this.h = addon.create(nx, ny, nz, this.dx, this.dz, this.pad); } catch (e) {
      this.key = null;
      if (/memory/.test(e.message) && !/^not enough GPU memory/.test(e.message)) throw new Error(`not enough GPU memory for this grid (it needs about ${(gridBytes(pr) / 1e9).toFixed(1)} GB): ${e.message}`);
      throw e;
    }
    GPU.track(this);
    this.U.set('pad', this.pad, nxp, nyp, 0);
    this.U.set('dims', nx, ny, nz, this.dx);
    this.U.set('dzv', this.dz, 1 / this.dx, 1 / this.dz, 0);
    const nv = (nx + 1) * (ny + 1) * (nz + 1) * 4, nc = nx * ny * nz * 4;
    this.V = [new Field(this, 'V', 0, nv), new Field(this, 'V', 1, nv)];
    this.Q1 = [new Field(this, 'Q1', 0, nc), new Field(this, 'Q1', 1, nc)];
    this.Q2 = [new Field(this, 'Q2', 0, nc), new Field(this, 'Q2', 1, nc)];
    this.Vt = new Field(this, 'Vt', 0, nv);
    this.aux = new Field(this, 'aux', 0, nx * ny * nz * 2);
    this.phi = new Field(this, 'phi', 0, nxp * nyp * nz);
    this.rhs = new Field(this, 'rhs', 0, nxp * nyp * nz);
    this.bs = new Field(this, 'bs', 0, (nz + 1) * 8);
    this.solver = { nx: nxp, ny: nyp, nz, dx: this.dx, dz: this.dz };
  }
  destroy() {
    if (!this.h) return;
    addon.destroy(this.h);
    GPU.untrack(this);
    this.h = null; this.V = null; this.solver = null;
  }
  // setSounding (model.js:1276-1288): the caller's Sounding module parses and analyses the text and builds the base
  // state; returns its analysis
  setSounding(text) {
    if (!SOUNDING) throw new Error('setSounding: no Sounding module (useSounding(S) first, or setBaseState)');
    const snd = SOUNDING.parseSharppy(text);
    const A = typeof SOUNDING.analyze === 'function' ? SOUNDING.analyze(snd) : null;
    this.analysis = A; this.snd = snd;
    this.setBaseState(SOUNDING.baseState(snd, this.nz, this.dz, [0, 0]));
    return A;
  }
  // the base state table (bsx.B: (nz + 1) * 8 floats, sounding.js:139-197) -> bs and the solver's densities
  setBaseState(bsx) {
    const B = Float32Array.from(bsx.B);
    if (B.length !== (this.nz + 1) * 8) throw new Error('setBaseState: ' + (this.nz + 1) * 8 + ' floats expected');
    this.base = Object.assign({}, bsx, { B }); this.rhoSfc = B[2];
    addon.setBaseState(this.h, B, this.nz);
  }
  // the extra wind field the trajectory-averaged pressure force needs (hurricane runs only)
  setPrepField(on) {
    addon.setPrepField(this.h, on ? 1 : 0);
    this.prep = !!on;
    this.Vp = on ? new Field(this, 'Vp', 0, (this.nx + 1) * (this.ny + 1) * (this.nz + 1) * 4) : null;
  }
  // hurricane mode. surf: Float32Array(nx*ny*4) = (land fraction, SST K, 0, 0) per column (null = no sea);
  // hur: {ck, landMoist, sfcP Pa, tauRad s, radCap K/s, blDepth m, u10f, gust m/s, ...} or null
  setSurface(surf, hur) {
    const a = surf ? Float32Array.from(surf) : (hur ? new Float32Array(this.nx * this.ny * 4) : null);
    if (a && a.length !== this.nx * this.ny * 4) throw new Error('setSurface: ' + this.nx * this.ny * 4 + ' floats expected');
    addon.setSurface(this.h, a);
    this.hur = hur || null;
    this.setPrepField(!!hur);
  }
  // the initial warm-core vortex: v = {x, y (model m), vmax m/s, rmw m, r0 m, ztop m, moistR m, moistRH, shape};
  // null = none. The (tangential wind, theta') table is built here and handed to the library for reset.
  setVortex(v) {
    if (!v) { this.vortex = null; addon.setVortex(this.h, null, 0); return null; }
    const r = vortexTable(this.nz, this.dz, this.dx, this.params.lat, this.base.B, v);
    addon.setVortex(this.h, r.tab, r.info.nr);
    this.vortex = r.info;
    return this.vortex;
  }
  baseWind(z) { const k = clamp(Math.floor(z / this.dz), 0, this.nz - 1), B = this.base.B; return [B[k * 8 + 6], B[k * 8 + 7]]; }
  setBoundaries(bnds) {
    this.bounds = bnds.slice(0, MAX_BOUNDS);
    const f = new Float32Array(MAX_BOUNDS * 148);
    this.bounds.forEach((b, i) => {
      const o = i * 148, pts = b.pts.slice(0, MAX_BPTS);
      f.set([pts.length, b.motion[0], b.motion[1], b.halfW, b.dT, b.Hmax, b.slope, b.Hmin, b.dq, b.qRatio, b.du, b.dv, b.wRatio, b.nudge ? 1 : 0, b.dFree, b.side,
        b.vary || 0, b.seed || 0, b.varyL || 60000, 0], o);
      pts.forEach((p, k) => { const v = (b.vel && b.vel[k]) || b.motion; f.set([p[0], p[1], v[0], v[1]], o + 20 + 4 * k); });
    });
    addon.setBoundaries(this.h, f, this.bounds.length);
  }
  setCI(events) {
    const act = events.filter(e => e.t1 > this.time).slice(-MAX_CI);
    this.ciActive = act.length;
    const f = new Float32Array(Math.max(1, act.length) * 12);
    act.forEach((e, i) => f.set([e.x, e.y, e.zc, e.rh, e.rz, e.wt, e.t0, e.t1, e.vx || 0, e.vy || 0, e.seed || i * 1.37, e.ramp || 150], i * 12));
    addon.setCI(this.h, f, act.length);
  }
  writeUniforms(dt) {
    const U = this.U, P = this.params;
    U.set('step', dt, this.time, this.rhoSfc || 1.15, 9.81);
    U.set('sponge', P.spongeW, P.spongeTau, P.dampZ, P.dampTau);
    U.set('turb', P.cs, P.nu0, P.evap, P.drag ? 1 : 0);
    const cd = Math.pow(0.41 / Math.log(0.5 * this.dz / P.z0 + 1), 2);
    U.set('micro', P.auto, P.qc0, P.freeze, cd);
    U.set('misc', P.vcap, P.mc, P.noise, P.ciRate);
    U.set('cor', 2 * 7.292e-5 * Math.sin(P.lat * DEG), P.nudgeTau, this.bounds.length, dflt(P.moistMix, 0));
    U.set('ci', this.ciActive || 0, 0, 0, 0);
    U.set('micro2', 0, dflt(P.sortQ, 5e-4), P.scheme === 'thompson' ? 1 : 0, dflt(P.evapT, 0.3));
    const H = this.hur;
    if (H) {
      U.set('hur', 1, H.ck, H.landMoist, H.sfcP);
      U.set('hur2', H.tauRad, H.radCap, H.blDepth, H.u10f);
      U.set('hur3', Math.pow(0.41 / Math.log(0.5 * this.dz / 0.3 + 1), 2), H.gust, dflt(H.thermals, 0) * dflt(H.thermalScale, 1), H.maritime ? 1 : 0);
      const f4 = dflt(H.filter4, 0.003);
      this.filtNow = this.steps % 4 === 3 && f4 > 0;
      U.set('hur5', dflt(H.cdCap, 2.4e-3), H.dissHeat ? 1 : 0, 4 * f4, 0);
      const C = H.centre || [this.Lx / 2, this.Ly / 2];
      U.set('hur4', C[0], C[1], H.thermalR || 200000, H.envR || 400000);
    } else {
      this.filtNow = false;
      U.set('hur', 0, 0, 0, 0); U.set('hur2', 1, 0, 1, 0.8); U.set('hur3', 0, 1, 0, 0); U.set('hur4', 0, 0, 1, 1e9); U.set('hur5', 2.4e-3, 0, 0, 0);
    }
    const cnOld = dflt(this.cnOld, CN_OLD), cn = !!H && !!this.prep && this.steps > 0 && cnOld > 0;
    this.cnNow = cn;
    U.set('cn', cn ? cnOld : 0, this.phiScale || 1, cn ? 1 / (1 - cnOld) : 1, H && dflt(this.traj2, true) ? 1 : 0);
    const X = this.vortex;
    if (X) { U.set('vtx', X.x, X.y, X.dr, X.nr); U.set('vtx2', X.moistR, X.moistRH, 1, 0); }
    else { U.set('vtx', 0, 0, 1, 0); U.set('vtx2', 1, 0, 0, 0); }
    U.set('dzv', this.dz, 1 / this.dx, 1 / this.dz, dflt(P.mcFloor, 0.25));
    return U.f;
  }
  reset() {
    this.time = 0; this.steps = 0; this.parity = 0; this.phiScale = 1;
    addon.reset(this.h, this.writeUniforms(1), this.noiseSeed || 0);
    this.lsfT = 0;
  }
  // build-defined warm bubble (no reference counterpart), on both state copies
  addBubble(x, y, z, rh, rz, dth) { addon.addBubble(this.h, x, y, z, rh, rz, dth); }
  // judge (with a surrogate attached): also run the plain microphysics from the same inputs
  encStep(enc, dt, judge) {
    void enc;
    const u = this.writeUniforms(dt);
    let flags = judge ? 2 : 0;
    if (this.time - this.lsfT >= 60 || this.time === 0) { flags |= 1; this.lsfT = this.time; }
    if (this.hur && this.filtNow) flags |= 4;  // the 2-dx filter first (hurricane mode, every 4th step)
    addon.step(this.h, u, this.parity, flags);
    this.parity = 1 - this.parity;
    this.time += dt; this.steps++;
    this.phiScale = this.cnNow ? 1 / (1 - dflt(this.cnOld, CN_OLD)) : 1;
  }
  setSolverMode(gemm, thomas, cond) { addon.setSolverMode(this.h, GEMM[gemm || 'bf16x6'], THOMAS[thomas || 'mixed'], dflt(cond, 64)); }
  sync() { return addon.sync(this.h); }
  // requestMaps (model.js:1478-1494): the 2-D maps of the current state from the uniform block last
  // written; returns () => Promise<{m0, m1, m2, m3: Float32Array(nx*ny*4), t}>, or null while both
  // readback slots are pending
  requestMaps(enc) {
    void enc;
    const slot = addon.mapsRequest(this.h, this.U.f, this.parity);
    if (slot < 0) return null;
    const t = this.time, n = this.nx * this.ny, h = this.h;
    return () => new Promise((resolve, reject) => setImmediate(() => {
      try {
        const a = new Float32Array(16 * n);
        addon.mapsCollect(h, slot, a);
        resolve({ m0: a.subarray(0, 4 * n), m1: a.subarray(4 * n, 8 * n), m2: a.subarray(8 * n, 12 * n), m3: a.subarray(12 * n, 16 * n), t });
      } catch (e) { reject(e); }
    }));
  }
  // ── microphysics surrogate (no reference counterpart) ──
  // teacher sampling of the microphysics block during plain steps (capacity 0: off)
  sampleMicro(capacity, pActive, pClear) { addon.sampleMicro(this.h, capacity, dflt(pActive, 1), dflt(pClear, 0)); }
  microSamples() { return addon.microSamples(this.h); }  // Float32Array, rows x 24
  // attach a trained Surrogate (null detaches); capacity 0 = half the cells
  setSurrogate(net, capacity) { this._net = net || null; addon.setSurrogate(this.h, net ? net.h : null, capacity || 0); }
  // usage and, over judged steps, the error against the plain block (relative L2 per channel, ...)
  surrogateStats(reset) {
    const s = addon.surrogateStats(this.h, !!reset);
    const rel = (d, r) => (r > 0 ? Math.sqrt(d / r) : (d === 0 ? 0 : Infinity));
    const names = ['qv', 'qc', 'qi', 'qr', 'qs', 'qg', 'Nr', 'theta'];
    const relL2 = {};
    names.forEach((n, i) => { relL2[n] = rel(s.d2[i], s.r2[i]); });
    return { steps: s.steps, points: s.points, fallbacks: s.fallbacks, overflow: s.overflow, judged: s.judged, relL2,
      columnCondensateRelL2: rel(s.ccD2, s.ccR2), dbzRms: s.dbzCols > 0 ? Math.sqrt(s.dbzD2 / s.dbzCols) : 0,
      dbzMaxAbsDiff: s.dbzMaxDiff, precipRelL2: rel(s.prD2, s.prR2), precipTotals: [s.prSumS, s.prSumP] };
  }
  get Vcur() { return this.V[this.parity]; }
}

// ── the microphysics surrogate network (no reference counterpart): 14 inputs -> 3 x 128 GELU -> 9
// outputs, trained on the GPU from the model's teacher rows (14 inputs, 9 tendencies, level) ──
class Surrogate {
  constructor() {
    this.h = addon.mlpCreate();
    this.n = addon.mlpNparams();
  }
  params() { const p = new Float32Array(this.n); addon.mlpParams(this.h, p, false); return p; }
  setParams(p) { addon.mlpParams(this.h, Float32Array.from(p), true); }
  setDeadZone(frac) { addon.mlpSetDeadZone(this.h, frac); }
  // the least total condensate (kg/kg) of a point handed to the network; thinner cloud runs the plain block
  setHandMin(q) { addon.mlpSetHandMin(this.h, q); }
  // the network's targets: the Nr output as log((Nr + dNr + 1) / (Nr + 1)); scales (RMS of each target)
  // and the input envelope (min / max) from the rows, as the Python host's `fit`
  static targets(rows) {
    const t = Float32Array.from(rows);
    for (let r = 0; r < t.length / 24; r++) {
      const n0 = Math.max(t[r * 24 + 9], 0), n1 = Math.max(n0 + t[r * 24 + 21], 0);
      t[r * 24 + 21] = Math.log((n1 + 1) / (n0 + 1));
    }
    return t;
  }
  fitScales(t) {
    const n = t.length / 24, s = new Float64Array(9), lo = new Float32Array(14).fill(Infinity), hi = new Float32Array(14).fill(-Infinity);
    for (let r = 0; r < n; r++) {
      for (let j = 0; j < 9; j++) { const v = t[r * 24 + 14 + j]; s[j] += v * v; }
      for (let i = 0; i < 14; i++) { const v = t[r * 24 + i]; if (v < lo[i]) lo[i] = v; if (v > hi[i]) hi[i] = v; }
    }
    const sc = Float32Array.from(s, v => (v > 0 ? Math.sqrt(v / n) : 1));
    addon.mlpSetScales(this.h, sc, lo, hi);
    return { scale: sc, lo, hi };
  }
  // GPU training on the rows (Adam; learning rate halved for each of the last three quarters)
  fit(rows, steps, opt) {
    const o = opt || {};
    const t = Surrogate.targets(rows);
    this.fitScales(t);
    addon.mlpLoadRows(this.h, t, 24);
    const chunk = o.chunk || 500, batch = o.batch || 4096, lr = o.lr || 2e-3, losses = [];
    for (let done = 0; done < steps; done += chunk) {
      const k = Math.min(chunk, steps - done);
      losses.push(addon.mlpTrain(this.h, k, batch, lr * Math.pow(0.5, Math.floor(4 * done / steps)), (o.seed || 1) + done, 1));
    }
    return losses;
  }
  infer(rows, stride) { return addon.mlpInfer(this.h, rows instanceof Float32Array ? rows : Float32Array.from(rows), stride || 24); }
}

// ── virtual WSR-88D radars: volume coverage patterns, moments and Archive II scales (radar.js:18-39) ──
const RE_M = 6371000, KE = 4 / 3;
const GATE0 = 2125, GATEDR = 250, MAX_REF_GATES = 1832, MAX_DOP_GATES = 1192;
const L2_SCALE = { REF: [2, 66], VEL: [2, 129], SW: [2, 129], ZDR: [16, 128], PHI: [253 / 360, 2], RHO: [300, -60.5] };
const CUT_MOMENTS = { surv: ['REF', 'ZDR', 'PHI', 'RHO'], dop: ['REF', 'VEL', 'SW'], both: ['REF', 'VEL', 'SW', 'ZDR', 'PHI', 'RHO'] };
const MOMENT_ORDER = ['REF', 'VEL', 'SW', 'ZDR', 'PHI', 'RHO'];
function splitVcp(split, rest, nyqSplit, nyqRest) {
  const cuts = [];
  split.forEach(a => {
    cuts.push({ angle: a, kind: 'surv', superRes: true, nyquist: 8.9 });
    cuts.push({ angle: a, kind: 'dop', superRes: true, nyquist: nyqSplit });
  });
  rest.forEach(a => cuts.push({ angle: a, kind: 'both', superRes: false, nyquist: a < 7 ? nyqRest : 32.3 }));
  return cuts;
}
const REST12 = [1.8, 2.4, 3.1, 4.0, 5.1, 6.4, 8.0, 10.0, 12.5, 15.6, 19.5];
const VCPS = {
  212: { label: 'VCP 212', dur: 270, cuts: splitVcp([0.5, 0.9, 1.3], REST12, 28.4, 29.4) },
  12: { label: 'VCP 12', dur: 250, cuts: splitVcp([0.5, 0.9, 1.3], REST12, 28.4, 29.4) },
  215: { label: 'VCP 215', dur: 360, cuts: splitVcp([0.5, 0.9, 1.3], [1.8, 2.4, 3.1, 4.0, 5.1, 6.4, 8.0, 10.0, 12.0, 14.0, 16.7, 19.5], 28.4, 29.4) },
  35: { label: 'VCP 35', dur: 420, cuts: splitVcp([0.5, 0.9, 1.3], [1.8, 2.4, 3.1, 4.0, 5.1, 6.4], 26.0, 27.0) },
};
const RADAR_FIELDS = ['site', 'geo', 'cut', 'gates', 'nyq', 'offA', 'offB', 'mdl', 'mdl2', 'phys', 'phys2'];

// the reference's seeded generator (rng.js:3-27): xmur3 of the seed's text, then mulberry32
class Rng {
  constructor(seed) {
    const str = String(seed);
    let h = 1779033703 ^ str.length;
    for (let i = 0; i < str.length; i++) {
      h = Math.imul(h ^ str.charCodeAt(i), 3432918353);
      h = (h << 13) | (h >>> 19);
    }
    h = Math.imul(h ^ (h >>> 16), 2246822507);
    h = Math.imul(h ^ (h >>> 13), 3266489909);
    this.state = (h ^ (h >>> 16)) >>> 0;
  }
  u() {
    this.state = (this.state + 0x6D2B79F5) >>> 0;
    let t = this.state;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  }
}

// Gate counts of a site: enough 250 m gates to reach the domain's farthest corner plus 10 km (at most 460 km),
// a multiple of 4 (four gates pack one 32-bit word of codes), the Doppler moments capped shorter (radar.js:289-300)
function gateCounts(sx, sy, Lx, Ly) {
  const far = Math.max(...[[0, 0], [Lx, 0], [0, Ly], [Lx, Ly]].map(c => Math.hypot(c[0] - sx, c[1] - sy)));
  const want = Math.ceil((Math.min(460000, far + 10000) - GATE0) / GATEDR);
  const ngRef = Math.max(64, Math.ceil(Math.min(MAX_REF_GATES, want) / 4) * 4);
  return { ngRef, ngDop: Math.min(MAX_DOP_GATES, ngRef) };
}
// A volume's byte layout and timing: cut by cut, each moment's radials x gates of 8-bit codes follow one another
// (the layout the Level II writer reads); each cut gets a share of the VCP's duration, super-resolution cuts
// twice the share of the others (radar.js:301-316)
function layVolume(vcp, ngRef, ngDop) {
  const share = c => (c.superRes ? 2 : 1);
  const shares = vcp.cuts.reduce((n, c) => n + share(c), 0);
  let bytes = 0, before = 0;
  const cuts = vcp.cuts.map((c, index) => {
    const nrad = c.superRes ? 720 : 360;
    const mom = CUT_MOMENTS[c.kind].map(name => {
      const ngates = name === 'REF' ? ngRef : ngDop, at = bytes;
      bytes += nrad * ngates;
      return { name, ngates, off: at };
    });
    const t0 = before / shares * vcp.dur;
    before += share(c);
    return Object.assign({}, c, { index, nrad, mom, t0, dur: share(c) / shares * vcp.dur });
  });
  return { cuts, bytes };
}

// Virtual WSR-88D radars over the model (the host half of radar.js's RadarSim, radar.js:271-418; the kernels are
// the library's). geo is the application's map projection (geo.js's GEO: toModel(lat, lon), lat0, lon0,
// groundMSL), as in the reference.
class RadarSim {
  constructor(model, geo) { this.m = model; this.geo = geo; this.radars = []; this.nTor = 0; }
  // sub-grid tornadoes: [{x, y, R, v, debrisTop, debrisDbz, debris}] model metres, at most 64
  setTornadoes(list) {
    const act = list.slice(0, 64), f = new Float32Array(Math.max(1, act.length) * 8);
    act.forEach((t, i) => f.set([t.x, t.y, t.R, t.v, t.debrisTop, t.debrisDbz, t.debris, 0], i * 8));
    addon.radarSetTornadoes(this.m.h, f, act.length);
    this.nTor = act.length;
  }
  // site: {id, lat, lon, h (ground m MSL)}; opts: {vcp, aliasing, minDbz, seed}
  addRadar(site, opts) {
    const R = RadarSim.plan(this.m, this.geo, site, opts);
    R.h = addon.radarCreate(this.m.h, R.bytes, 720 * R.ngDop);
    this.radars.push(R);
    return R;
  }
  // a radar's layout, timing and schedule state, without its device buffers
  static plan(m, geo, site, opts) {
    const o = opts || {};
    const vcp = VCPS[o.vcp] || VCPS[212];
    const [sx, sy] = geo.toModel(site.lat, site.lon);
    const { ngRef, ngDop } = gateCounts(sx, sy, m.Lx, m.Ly);
    const { cuts, bytes } = layVolume(vcp, ngRef, ngDop);
    const rng = o.seed !== undefined ? new Rng(o.seed) : null;
    const rand = rng ? () => rng.u() : Math.random;
    return {
      site, opts: o, vcp, vcpId: +o.vcp || 212, cuts, bytes, ngRef, ngDop, sx, sy, rand, U: new UBlock(RADAR_FIELDS), h: null,
      readBusy: [false, false], phase: rand() * vcp.dur, vol: -1, next: 0, volStart: 0, az0: 0,
    };
  }
  destroy() {
    if (this.m.h) this.radars.forEach(R => addon.radarDestroy(R.h));  // else the model freed them
    this.radars = [];
  }
  // Enqueue, in scan order, every cut whose start the model clock simT has reached; returns the volumes that
  // completed (their codes are on their way to a readback slot).
  encDue(enc, simT) {
    const finished = [];
    for (const R of this.radars) {
      for (let c = RadarSim.dueCut(R, simT); c; c = RadarSim.dueCut(R, simT)) {
        this.encCut(enc, R, c, simT);
        if (++R.next === R.cuts.length) finished.push(this.endVolume(R));
      }
    }
    return finished;
  }
  // the radar's next cut if simT has reached it, opening the next volume (a new random start azimuth) once the
  // last one is complete, its start time has come and a readback slot is free for it; else null
  static dueCut(R, simT) {
    if (R.vol < 0 || R.next >= R.cuts.length) {
      const start = R.vol < 0 ? R.phase : R.volStart + R.vcp.dur;
      if (simT < start || R.readBusy.every(Boolean)) return null;
      R.vol += 1; R.volStart = start; R.next = 0; R.az0 = R.rand();
    }
    const c = R.cuts[R.next];
    return simT >= R.volStart + c.t0 ? c : null;
  }
  // the complete volume goes to a readback slot
  endVolume(R) {
    const slot = addon.radarReadback(this.m.h, R.h);
    if (slot < 0) throw new Error('radar ' + R.site.id + ': both readback slots are pending');
    R.readBusy[slot] = true;
    return { R, slot, vol: R.vol, volStart: R.volStart, az0: R.az0 };
  }
  cutBlock(R, c, simT) { return RadarSim.block(this.m, this.geo, R, c, simT, this.nTor); }
  // the cut's RADAR_FIELDS block (radar.js:369-384): site, projection, cut geometry, gates, Nyquist and clock, each
  // moment's word offset in the volume with a mask of the moments present, the model grid and its physics switches
  static block(m, geo, R, c, simT, nTor) {
    const U = R.U, word = {};
    c.mom.forEach(x => { word[x.name] = x.off / 4; });
    const mask = MOMENT_ORDER.reduce((b, n, i) => (n in word ? b | (1 << i) : b), 0);
    const w = n => (n in word ? word[n] : 0);
    U.set('site', R.sx, R.sy, R.site.h + 20, KE * RE_M);
    U.set('geo', geo.lat0 * DEG, geo.lon0 * DEG, R.site.lat * DEG, R.site.lon * DEG);
    U.set('cut', c.angle * DEG, R.az0, 360 / c.nrad, c.nrad);
    U.set('gates', R.ngRef, R.ngDop, GATE0, GATEDR);
    U.set('nyq', R.opts.aliasing === false ? 63.5 : c.nyquist, simT, (R.vol * 97 + c.index * 13) % 100000, 25);
    U.set('offA', w('REF'), w('VEL'), w('SW'), w('ZDR'));
    U.set('offB', w('PHI'), w('RHO'), 0, mask);
    U.set('mdl', m.nx, m.ny, m.nz, m.dx);
    U.set('mdl2', m.dz, geo.groundMSL || 0, -m.Lx / 2, -m.Ly / 2);
    U.set('phys', m.rhoSfc || 1.15, dflt(R.opts.minDbz, 0), m.params.scheme === 'thompson' ? 1 : 0, nTor || 0);
    U.set('phys2', m.hur && m.hur.maritime ? 1 : 0, 0, 0, 0);
    return U.f;
  }
  encCut(enc, R, c, simT) {
    void enc;
    addon.radarCut(this.m.h, R.h, this.cutBlock(R, c, simT), this.m.parity);
    c.simT = simT;
  }
  // the header the upload server reads for a finished volume (radar.js:394-416; server.py:249-284)
  static header(R, d, run, epoch0) {
    const t0 = dflt(epoch0, 0) + d.volStart * 1000, nyq = c => (R.opts.aliasing === false ? 63.5 : c.nyquist);
    const sweep = c => ({
      angle: c.angle, superRes: c.superRes, kind: c.kind, az0: d.az0, nrad: c.nrad, nyquist: nyq(c),
      unamb_km: c.kind === 'surv' ? 466 : 300, start_ms: Math.round(t0 + c.t0 * 1000), dur_ms: Math.round(c.dur * 1000),
      moments: c.mom.map(x => ({ name: x.name, ngates: x.ngates, off: x.off })),
    });
    return { run, icao: R.site.id, lat: R.site.lat, lon: R.site.lon, elev: R.site.h + 20, vcp: R.vcpId, epoch_ms: Math.round(t0), sweeps: R.cuts.map(sweep) };
  }
  // a finished volume's codes and header; frees its readback slot
  collect(d, run, epoch0) {
    const R = d.R;
    return new Promise((resolve, reject) => setImmediate(() => {
      try {
        const bytes = new Uint8Array(R.bytes);
        addon.radarCollect(this.m.h, R.h, d.slot, bytes);  // a failed collect leaves the slot pending
        R.readBusy[d.slot] = false;
        resolve({ R, header: RadarSim.header(R, d, run, epoch0), bytes });
      } catch (e) { reject(e); }
    }));
  }
}

// the /api/volume request body of a collected volume (app.js:1743-1750): u32-LE JSON length, the header's
// JSON, the codes; gzip (sent with X-OB-Gzip: 1)
function volumeBody(header, bytes) {
  const hj = Buffer.from(JSON.stringify(header), 'utf8');
  const pre = Buffer.alloc(4);
  pre.writeUInt32LE(hj.length, 0);
  return zlib.gzipSync(Buffer.concat([pre, hj, Buffer.from(bytes.buffer, bytes.byteOffset, bytes.byteLength)]));
}

module.exports = { vortexTable, volumeBody, RadarSim, Rng, VCPS, L2_SCALE, RADAR_FIELDS, Surrogate, RegionalModel, PoissonDirect, useSounding, UBlock, MODEL_FIELDS, GPU, levelsFor, customGrid, gridPad, gridBytes, GRID_PRESETS, addon };
