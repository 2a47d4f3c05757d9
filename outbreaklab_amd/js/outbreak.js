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
const lerp = (a, b, t) => a + (b - a) * t;
const dflt = (v, d) => (v === undefined || v === null ? d : v);
const FIELD = { V: 0, Q1: 1, Q2: 2, Vt: 3, Q1t: 4, Q2t: 5, aux: 6, phT: 7, lsfA: 8, lsfB: 9, ciMap: 10, phi: 11, rhs: 12, bs: 13, sfc: 14, Vp: 15 };
const GEMM = { fp32: 0, bf16x3: 1 };
const THOMAS = { fp32: 0, fp64: 1, mixed: 2 };
const MAX_BOUNDS = 8, MAX_BPTS = 32, MAX_CI = 256, VX_MAXR = 2048;
const CN_OLD = 0.4;  // hurricane mode: the share of the last step's pressure force applied before advection
const smoothstep = (a, b, x) => { const t = clamp((x - a) / (b - a), 0, 1); return t * t * (3 - 2 * t); };

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

// ── grid helpers (model.js:1104-1154) ──
const LEVELS_BY_DX = [[100, 160], [250, 100], [500, 80], [1000, 64], [1500, 60], [2000, 56], [3000, 52], [4000, 48]];
function levelsFor(dx) {
  const T = LEVELS_BY_DX, l = Math.log(clamp(dx, T[0][0], T[T.length - 1][0]));
  for (let i = 0; i < T.length - 1; i++) {
    const a = Math.log(T[i][0]), b = Math.log(T[i + 1][0]);
    if (l <= b) return Math.round(lerp(T[i][1], T[i + 1][1], (l - a) / (b - a)) / 4) * 4;
  }
  return T[T.length - 1][1];
}
function customGrid(c) {
  const dx = clamp(+c.dx || 500, 100, 4000);
  const side = km => clamp(Math.round((+km || 256) * 1000 / dx / 8) * 8, 64, 2040);
  const nx = side(c.w), ny = side(c.h), nz = levelsFor(dx);
  return { key: `custom:${dx}:${nx}x${ny}x${nz}`, dx, n: [nx, ny, nz] };
}
function gridPad(dx) { return clamp(Math.round(10000 / dx / 8) * 8, 8, 64); }
function gridBytes(g) {
  const [nx, ny, nz] = g.n, p = gridPad(g.dx), nxp = nx + 2 * p, nyp = ny + 2 * p;
  return 1.05 * ((nx + 1) * (ny + 1) * (nz + 1) * 48 + nx * ny * nz * 108 + nxp * nyp * nz * 20 + nx * ny * 192 + 8 * (nxp * nxp + nyp * nyp));
}
const GRID_PRESETS = {
  quick: { dx: 2000, n: [320, 320, 56] }, standard: { dx: 1500, n: [448, 448, 60] }, large: { dx: 1500, n: [640, 640, 60] },
  fine: { dx: 1000, n: [640, 640, 64] }, m500: { dx: 500, n: [512, 512, 80] }, m250: { dx: 250, n: [512, 512, 100] },
  m100: { dx: 100, n: [512, 512, 160] }, h2: { dx: 2000, n: [640, 640, 56] }, h3: { dx: 3000, n: [640, 640, 52] },
  h15: { dx: 1500, n: [960, 960, 60] },
};

// ── soundings: SHARPpy parser, winds and the hydrostatic base state (sounding.js:10-197) ──
const Sounding = (function () {
  const KT = 0.514444, RD = 287.04, CP = 1005.7, EPS = 0.622, G = 9.80665;
  function parseSharppy(text) {
    const out = { title: '', station: '', time: null, levels: [] };
    let mode = '';
    for (const raw of text.replace(/\r/g, '').split('\n')) {
      const L = raw.trim();
      if (!L) continue;
      if (/^%TITLE%/i.test(L)) { mode = 'title'; continue; }
      if (/^%RAW%/i.test(L)) { mode = 'raw'; continue; }
      if (/^%END%/i.test(L)) { mode = ''; continue; }
      if (mode === 'title') { if (!out.title) { out.title = L; out.station = L.split(/\s+/)[0]; } continue; }
      if (mode === 'raw' || (!mode && /^-?\d/.test(L))) {
        const f = L.split(/[,\s]+/).filter(Boolean).map(Number);
        if (f.length < 6 || f.some(v => !isFinite(v))) continue;
        const [p, h, t, td, wd, ws] = f;
        if (p <= 0 || h < -1000 || h > 40000 || p === -9999) continue;
        out.levels.push({ p, h, t: t > -9990 ? t : NaN, td: td > -9990 ? td : NaN, wd: wd > -9990 ? wd : NaN, ws: ws > -9990 ? ws : NaN });
      }
    }
    out.levels.sort((a, b) => b.p - a.p);
    if (out.levels.length < 4) throw new Error('Could not read at least 4 levels. Expected SHARPpy format: %RAW% rows of  PRES, HGHT, TEMP, DWPT, WDIR, WSPD(kt)  ... %END%');
    return out;
  }
  function es(t) { return 6.112 * Math.exp(17.67 * t / (t + 243.5)); }
  function mixr(p, t) { const e = es(t); return EPS * e / Math.max(p - e, 1e-3); }
  function winds(snd) {
    const lv = snd.levels.filter(l => isFinite(l.h));
    const sfcH = lv[0].h;
    const wl = lv.filter(l => isFinite(l.wd) && isFinite(l.ws)).map(l => {
      const d = l.wd * Math.PI / 180, s = l.ws * KT;
      return { z: (l.h - sfcH) / 1000, u: -s * Math.sin(d), v: -s * Math.cos(d) };
    });
    return function windAt(z) {
      if (z <= wl[0].z) return [wl[0].u, wl[0].v];
      for (let i = 0; i < wl.length - 1; i++) {
        if (z >= wl[i].z && z <= wl[i + 1].z) {
          const f = (z - wl[i].z) / Math.max(1e-6, wl[i + 1].z - wl[i].z);
          return [wl[i].u + (wl[i + 1].u - wl[i].u) * f, wl[i].v + (wl[i + 1].v - wl[i].v) * f];
        }
      }
      const l = wl[wl.length - 1];
      return [l.u, l.v];
    };
  }
  function baseState(snd, nz, dz, motion) {
    const lv = snd.levels.filter(l => isFinite(l.h) && isFinite(l.t)).slice().sort((a, b) => a.h - b.h);
    const sfcH = lv[0].h, sfcP = lv[0].p, top = lv[lv.length - 1];
    const byH = (key, zAgl) => {
      const h = zAgl + sfcH;
      if (h <= lv[0].h) return lv[0][key];
      for (let i = 0; i < lv.length - 1; i++) {
        if (h <= lv[i + 1].h) {
          const a = lv[i], b = lv[i + 1], f = (h - a.h) / Math.max(b.h - a.h, 1e-3), va = a[key], vb = b[key];
          if (!isFinite(va) || !isFinite(vb)) return isFinite(va) ? va : vb;
          if (key === 'p') return Math.exp(Math.log(va) + (Math.log(vb) - Math.log(va)) * f);
          return va + (vb - va) * f;
        }
      }
      return NaN;
    };
    const windAt = winds(snd);
    const B = new Float32Array((nz + 1) * 8), th = [], qv = [];
    for (let k = 0; k < nz; k++) {
      const z = (k + 0.5) * dz;
      let T = byH('t', z), Td = byH('td', z), p = byH('p', z);
      if (!isFinite(T) || !isFinite(p)) {
        const dzt = z + sfcH - top.h;
        T = top.t; p = top.p * Math.exp(-G * dzt / (RD * (top.t + 273.15))); Td = Math.min(top.td, -80);
      }
      if (!isFinite(Td)) Td = T - 30;
      Td = Math.min(Td, T - 0.3);
      th.push((T + 273.15) * Math.pow(1000 / p, RD / CP));
      qv.push(Math.max(1e-7, mixr(p, Td)));
    }
    const piF = new Float64Array(nz + 1);
    piF[0] = Math.pow(sfcP / 1000, RD / CP);
    for (let k = 0; k < nz; k++) piF[k + 1] = piF[k] - G * dz / (CP * th[k] * (1 + 0.61 * qv[k]));
    for (let k = 0; k < nz; k++) {
      const pi = 0.5 * (piF[k] + piF[k + 1]), p = 100000 * Math.pow(pi, CP / RD), T = th[k] * pi;
      const esx = 611.2 * Math.exp(17.67 * (T - 273.15) / (T - 29.65)), qvs = EPS * esx / (p - esx);
      qv[k] = Math.min(qv[k], 0.97 * qvs);
      const rho = p / (RD * T * (1 + 0.61 * qv[k])), w = windAt((k + 0.5) * dz / 1000);
      B.set([th[k], qv[k], rho, 0, pi, p, w[0] - motion[0], w[1] - motion[1]], k * 8);
    }
    for (let k = 0; k <= nz; k++) {
      B[k * 8 + 3] = k === 0 ? B[2] + (B[2] - B[10]) * 0.5 : k === nz ? B[(nz - 1) * 8 + 2] * 0.94 : 0.5 * (B[(k - 1) * 8 + 2] + B[k * 8 + 2]);
    }
    return { B, sfcP, sfcH };
  }
  return { parseSharppy, baseState, winds, mixr, KT };
})();

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
  // host arrays (x fastest); precision: {gemm: 'fp32'|'bf16x3', thomas: 'fp32'|'fp64'|'mixed', cond}
  solve(rhs, opt) {
    const o = opt || {};
    const phi = new Float32Array(rhs.length);
    addon.poissonSolve(this.h, rhs instanceof Float32Array ? rhs : Float32Array.from(rhs), phi,
      GEMM[o.gemm || 'fp32'], THOMAS[o.thomas || 'mixed'], dflt(o.cond, 16));
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

// setVortex's table (model.js:1318-1365): v(r, z) = Vs(r) F(z), Vs a Rankine-type profile made inertially stable,
// phi from the gradient-wind balance, theta' = theta0 (d phi / dz) / g; the same arithmetic as the Python host
function vortexTable(nz, dz, dx, lat, B, v) {
  const f = 2 * 7.292e-5 * Math.sin(lat * DEG), dr = Math.max(500, dx / 2), nr = Math.min(VX_MAXR, Math.ceil(v.r0 * 2 / dr) + 3);
  const sg = f < 0 ? -1 : 1, fa = Math.abs(f), al = 0.6, vs = new Float64Array(nr);
  let Mp = 0;
  for (let i = 0; i < nr; i++) {
    const r = i * dr, x = r / v.rmw;
    let s = v.vmax * (1 + al) * x / (al + Math.pow(x, 1 + al)) * (1 - smoothstep(0.5 * v.r0, v.r0, r));
    if (r > 0) s = Math.max(s, (Mp - 0.5 * fa * r * r) / r, 0);
    vs[i] = s; Mp = r * s + 0.5 * fa * r * r;
  }
  const Vs = r => sg * vs[Math.min(nr - 1, Math.round(r / dr))];
  const Fz = v.shape === 'mid' ? z => Math.exp(-Math.pow((z - 5000) / 2500, 2))
    : z => (z >= v.ztop ? 0 : 1 - (z / v.ztop - Math.sin(2 * Math.PI * z / v.ztop) / (2 * Math.PI)));
  const vt = [], ph = [];
  for (let k = 0; k < nz; k++) {
    const F = Fz((k + 0.5) * dz), vk = new Float64Array(nr), pk = new Float64Array(nr);
    for (let i = 0; i < nr; i++) vk[i] = Vs(i * dr) * F;
    const g = i => (i === 0 ? 0 : vk[i] * vk[i] / (i * dr) + f * vk[i]);
    for (let i = nr - 2; i >= 0; i--) pk[i] = pk[i + 1] - 0.5 * (g(i) + g(i + 1)) * dr;
    vt.push(vk); ph.push(pk);
  }
  const tab = new Float32Array(nz * nr * 4);
  for (let k = 0; k < nz; k++) for (let i = 0; i < nr; i++) {
    const k0 = Math.max(k - 1, 0), k1 = Math.min(k + 1, nz - 1), b = (ph[k1][i] - ph[k0][i]) / ((k1 - k0) * dz);
    tab[(k * nr + i) * 4] = vt[k][i]; tab[(k * nr + i) * 4 + 1] = B[k * 8] * b / 9.81;
  }
  let minStab = 1e9, warm = 0;
  for (let k = 1; k < nz; k++) {
    const z = k * dz;
    if (z < 2000 || z > v.ztop) continue;
    for (let i = 0; i < Math.min(nr, Math.ceil(2 * v.rmw / dr)); i++) {
      const a = B[(k - 1) * 8] + tab[((k - 1) * nr + i) * 4 + 1], b = B[k * 8] + tab[(k * nr + i) * 4 + 1];
      minStab = Math.min(minStab, (b - a) / dz * 1000);
    }
    warm = Math.max(warm, tab[(k * nr) * 4 + 1]);
  }
  const info = { x: v.x, y: v.y, dr, nr, moistR: v.moistR || 1, moistRH: v.moistR ? v.moistRH : 0, pc: ph[0][0] * B[2] / 100, minStab, warm };
  return { tab, info };
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
  setSounding(text) {
    const snd = Sounding.parseSharppy(text);
    const bsx = Sounding.baseState(snd, this.nz, this.dz, [0, 0]);
    this.snd = snd; this.base = bsx; this.rhoSfc = bsx.B[2];
    addon.setBaseState(this.h, bsx.B, this.nz);
    return { snd, windAt: Sounding.winds(snd) };
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
  setSolverMode(gemm, thomas, cond) { addon.setSolverMode(this.h, GEMM[gemm || 'fp32'], THOMAS[thomas || 'mixed'], dflt(cond, 16)); }
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

// ── virtual WSR-88D radars (radar.js:18-418), azimuthal-equidistant projection (geo.js:6-29) ──
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

class Geo {
  constructor(lat0, lon0, Lx, Ly, groundMSL) {
    Object.assign(this, { lat0, lon0, Lx: Lx || 0, Ly: Ly || 0, groundMSL: groundMSL || 0 });
    this.p0 = lat0 * DEG; this.l0 = lon0 * DEG; this.sp0 = Math.sin(this.p0); this.cp0 = Math.cos(this.p0);
  }
  fwd(lat, lon) {
    const p = lat * DEG, dl = lon * DEG - this.l0, sp = Math.sin(p), cp = Math.cos(p), cdl = Math.cos(dl);
    const c = Math.acos(clamp(this.sp0 * sp + this.cp0 * cp * cdl, -1, 1));
    const k = c < 1e-9 ? 1 : c / Math.sin(c);
    return [RE_M * k * cp * Math.sin(dl), RE_M * k * (this.cp0 * sp - this.sp0 * cp * cdl)];
  }
  inv(x, y) {
    const rho = Math.hypot(x, y);
    if (rho < 1e-6) return [this.lat0, this.lon0];
    const c = rho / RE_M, sc = Math.sin(c), cc = Math.cos(c);
    const p = Math.asin(cc * this.sp0 + y * sc * this.cp0 / rho);
    const l = this.l0 + Math.atan2(x * sc, rho * this.cp0 * cc - y * this.sp0 * sc);
    return [p / DEG, ((l / DEG + 540) % 360) - 180];
  }
  toModel(lat, lon) { const xy = this.fwd(lat, lon); return [xy[0] + this.Lx / 2, xy[1] + this.Ly / 2]; }
  fromModel(mx, my) { return this.inv(mx - this.Lx / 2, my - this.Ly / 2); }
}

class RadarSim {
  constructor(model, geo) { this.m = model; this.geo = geo; this.radars = []; this.nTor = 0; }
  // tornadoes: [{x, y, R, v, debrisTop, debrisDbz, debris}] model metres
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
  // addRadar's layout without the device buffers (radar.js:289-316)
  static plan(m, g, site, opts) {
    opts = opts || {};
    const vcp = VCPS[opts.vcp] || VCPS[212];
    const sxy = g.toModel(site.lat, site.lon), sx = sxy[0], sy = sxy[1];
    let far = 0;
    [[0, 0], [m.Lx, 0], [0, m.Ly], [m.Lx, m.Ly]].forEach(c => { far = Math.max(far, Math.hypot(c[0] - sx, c[1] - sy)); });
    const rng = Math.min(460000, far + 10000);
    let ngRef = Math.min(MAX_REF_GATES, Math.ceil((rng - GATE0) / GATEDR));
    ngRef = Math.max(64, Math.ceil(ngRef / 4) * 4);
    const ngDop = Math.min(MAX_DOP_GATES, ngRef);
    let off = 0;
    const cuts = vcp.cuts.map((c, ci) => {
      const nrad = c.superRes ? 720 : 360;
      const mom = CUT_MOMENTS[c.kind].map(name => {
        const ng = name === 'REF' ? ngRef : ngDop, o = { name, ngates: ng, off };
        off += nrad * ng;
        return o;
      });
      return Object.assign({}, c, { index: ci, nrad, mom });
    });
    const wsum = cuts.reduce((s, c) => s + (c.superRes ? 2 : 1), 0);
    let acc = 0;
    cuts.forEach(c => { const w = c.superRes ? 2 : 1; c.t0 = acc / wsum * vcp.dur; acc += w; c.dur = w / wsum * vcp.dur; });
    const rand = opts.seed !== undefined ? (r => () => r.u())(new Rng(opts.seed)) : Math.random;
    return {
      site, opts, vcp, vcpId: +opts.vcp || 212, cuts, bytes: off, ngRef, ngDop, sx, sy, rand, U: new UBlock(RADAR_FIELDS),
      h: null, readBusy: [false, false], phase: rand() * vcp.dur, vol: -1, next: 0, volStart: 0, cutsDone: 0, az0: 0,
      done: 0, failed: 0,
    };
  }
  destroy() {
    if (this.m.h) this.radars.forEach(R => addon.radarDestroy(R.h));  // else the model freed them
    this.radars = [];
  }
  // encode every cut whose scan time has been reached; returns the volumes that completed
  encDue(enc, simT) {
    const done = [];
    for (const R of this.radars) {
      for (;;) {
        if (R.vol < 0 || R.next >= R.cuts.length) {
          const vstart = R.vol < 0 ? R.phase : R.volStart + R.vcp.dur;
          if (simT < vstart || (R.readBusy[0] && R.readBusy[1])) break;
          R.vol++; R.volStart = vstart; R.next = 0; R.az0 = R.rand();
        }
        const c = R.cuts[R.next];
        if (simT < R.volStart + c.t0) break;
        this.encCut(enc, R, c, simT);
        R.next++;
        if (R.next >= R.cuts.length) {
          const slot = addon.radarReadback(this.m.h, R.h);
          if (slot < 0) throw new Error('radar ' + R.site.id + ': both readback slots are pending');
          R.readBusy[slot] = true;
          done.push({ R, slot, vol: R.vol, volStart: R.volStart, az0: R.az0 });
        }
      }
    }
    return done;
  }
  // the cut's RADAR_FIELDS block (radar.js:369-382)
  cutBlock(R, c, simT) { return RadarSim.block(this.m, this.geo, R, c, simT, this.nTor); }
  static block(m, g, R, c, simT, nTor) {
    const U = R.U, off = {};
    c.mom.forEach(o => { off[o.name] = o.off / 4; });
    let mask = 0;
    MOMENT_ORDER.forEach((n, i) => { if (off[n] !== undefined) mask |= 1 << i; });
    const o = n => dflt(off[n], 0);
    U.set('site', R.sx, R.sy, R.site.h + 20, KE * RE_M);
    U.set('geo', g.lat0 * DEG, g.lon0 * DEG, R.site.lat * DEG, R.site.lon * DEG);
    U.set('cut', c.angle * DEG, R.az0, 360 / c.nrad, c.nrad);
    U.set('gates', R.ngRef, R.ngDop, GATE0, GATEDR);
    U.set('nyq', R.opts.aliasing === false ? 63.5 : c.nyquist, simT, (R.vol * 97 + c.index * 13) % 100000, 25);
    U.set('offA', o('REF'), o('VEL'), o('SW'), o('ZDR'));
    U.set('offB', o('PHI'), o('RHO'), 0, mask);
    U.set('mdl', m.nx, m.ny, m.nz, m.dx);
    U.set('mdl2', m.dz, g.groundMSL, -m.Lx / 2, -m.Ly / 2);
    U.set('phys', m.rhoSfc || 1.15, dflt(R.opts.minDbz, 0), m.params.scheme === 'thompson' ? 1 : 0, nTor || 0);
    U.set('phys2', m.hur && m.hur.maritime ? 1 : 0, 0, 0, 0);  // radar.js:384
    return U.f;
  }
  encCut(enc, R, c, simT) {
    void enc;
    addon.radarCut(this.m.h, R.h, this.cutBlock(R, c, simT), this.m.parity);
    c.simT = simT;
  }
  // a finished volume's bytes and the header the server reads (radar.js:394-416)
  collect(d, run, epoch0) {
    const R = d.R;
    return new Promise((resolve, reject) => setImmediate(() => {
      try {
        const bytes = new Uint8Array(R.bytes);
        addon.radarCollect(this.m.h, R.h, d.slot, bytes);  // a failed collect leaves the slot pending
        R.readBusy[d.slot] = false;
        const t0 = dflt(epoch0, 0) + d.volStart * 1000;
        const header = {
          run, icao: R.site.id, lat: R.site.lat, lon: R.site.lon, elev: R.site.h + 20, vcp: R.vcpId, epoch_ms: Math.round(t0),
          sweeps: R.cuts.map(c => ({
            angle: c.angle, superRes: c.superRes, kind: c.kind, az0: d.az0, nrad: c.nrad,
            nyquist: R.opts.aliasing === false ? 63.5 : c.nyquist, unamb_km: c.kind === 'surv' ? 466 : 300,
            start_ms: Math.round(t0 + c.t0 * 1000), dur_ms: Math.round(c.dur * 1000),
            moments: c.mom.map(o => ({ name: o.name, ngates: o.ngates, off: o.off })),
          })),
        };
        resolve({ R, header, bytes });
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

module.exports = { vortexTable, volumeBody, RadarSim, Geo, Rng, VCPS, L2_SCALE, RADAR_FIELDS, Surrogate, RegionalModel, PoissonDirect, Sounding, UBlock, MODEL_FIELDS, GPU, levelsFor, customGrid, gridPad, gridBytes, GRID_PRESETS, addon };
