/*
 * NSLab on MI355X — Node host module. A drop-in for pocket-wind-tunnel/src/nslab.js: the same exports
 * ({ VERSION, IC_INFO, createSolver, studySummary, Spectral, fftPlan, fft1, factorize }, nslab.js:668) and the same
 * solver object (nslab.js:648), with every transform, right-hand side, RK4 step and diagnostic computed in FP64 on the
 * GPU by libnslab_hip.so through the N-API addon. This file keeps the host-side bookkeeping the reference keeps on
 * the host: the mutable `st` object and its series (nslab.js:188-190, 300-302); the health report, the dossier and
 * the refinement verdict (nslab.js:595-666) are built by report.js from the numbers the GPU produced.
 *
 * There is no CPU path: if the addon cannot be loaded the require() throws.
 */
'use strict';
const path = require('path');
const addon = require(path.join(__dirname, '..', 'lib', 'nslab_napi.node'));
const report = require('./report.js');

const VERSION = '0.1.2';
const IC_KIND = { tgv: 0, tgv2d: 1, abc: 2, tubes: 3, random: 4, tubesz: 5 };

const IC_INFO = {
  tgv: { name: 'Taylor–Green vortex (3D)', exact: null, desc: 'u = sin x cos y cos z, v = −cos x sin y cos z, w = 0 (Brachet et al. 1983)' },
  tgv2d: { name: 'Taylor–Green vortex (2D, exact decay)', exact: 'e^{-2νt}', desc: 'u = sin x cos y, v = −cos x sin y, w = 0 — exact solution ∝ e^{−2νt}' },
  abc: { name: 'Arnold–Beltrami–Childress (exact decay)', exact: 'e^{-νt}', desc: 'u = A sin z + C cos y, v = B sin x + A cos z, w = C sin y + B cos x — Beltrami, exact solution ∝ e^{−νt}' },
  tubes: { name: 'Antiparallel vortex tubes', exact: null, desc: 'two Gaussian-core tubes of opposite circulation along x, sinusoidally perturbed (Kerr-type reconnection / stretching experiment)' },
  random: { name: 'Random solenoidal field', exact: null, desc: 'Gaussian random field with E(k) ∝ k⁴ exp(−2(k/k₀)²), seeded (decaying isotropic turbulence)' },
  // not in the reference: the tube pair of the memory-heavy benchmark (Gaussian cores along z, velocity from vorticity)
  tubesz: { name: 'Antiparallel vortex tubes along z', exact: null, desc: 'circulation ±Γ, core radius σ, centrelines (π ± (d + 0.1 cos z), π), plus a seeded solenoidal perturbation of rms velocity pert' },
};

// IC parameter vector with the reference's defaults (nslab.js:255-274).
function icVector(kind, prm) {
  prm = prm || {};
  const nn = (v, d) => (v != null ? +v : d);
  if (!(kind in IC_KIND)) throw new Error('unknown initial condition ' + kind);
  if (kind === 'abc') return [nn(prm.A, 1), nn(prm.B, 1), nn(prm.C, 1), 0];
  if (kind === 'tubes') return [nn(prm.amp, 8), nn(prm.sigma, 0.4), nn(prm.sep, 0.7), nn(prm.pert, 0.2)];
  if (kind === 'tubesz') return [nn(prm.Gamma, 1), nn(prm.sigma, 0.3), nn(prm.sep, 0.8), nn(prm.pert, 1e-4)];
  if (kind === 'random') return [(prm.seed || 1) | 0, prm.k0 || 4, prm.E0 || 0.5, 0];
  return [0, 0, 0, 0];
}

// ------------------------------------------------------------------------------------------------ FFT exports
function factorize(N) { return addon.factorize(N); }

/** Plan object with the reference's shape (nslab.js:36-44): N, stages, and the line buffers fft1 transforms. */
function fftPlan(N) {
  const f = factorize(N), stages = []; let n = N;
  for (const r of f) { stages.push({ r, n, m: n / r }); n /= r; }
  return { N, stages, xr: new Float64Array(N), xi: new Float64Array(N), yr: new Float64Array(N), yi: new Float64Array(N) };
}
/** Transform of plan.xr/xi on the GPU; sign −1 forward, +1 inverse, unnormalised. Returns [re, im] (nslab.js:46). */
function fft1(plan, sign) {
  const re = Float64Array.from(plan.xr), im = Float64Array.from(plan.xi);
  addon.fft1(plan.N, sign, re, im);
  plan.yr.set(re); plan.yi.set(im);
  return [plan.yr, plan.yi];
}

/** 3D real <-> half-spectrum transforms (nslab.js:94-155); layout (kz*N + ky)*NH + kx. */
class Spectral {
  constructor(N) {
    factorize(N);
    this.N = N; this.NH = N / 2 + 1; this.NS = N * N * this.NH; this.N3 = N * N * N;
    this.plan = fftPlan(N);
  }
  forward(f, re, im) { addon.spectralForward(this.N, f, re, im); }
  inverse(re, im, f) { addon.spectralInverse(this.N, re, im, f); }
}

// ------------------------------------------------------------------------------------------------ solver
function createSolver(opts, pool) {
  const N = opts.N || 32, NH = N / 2 + 1, NS = N * N * NH, N3 = N * N * N;
  const nu = opts.nu != null ? opts.nu : 1 / (opts.Re || 100), Re = 1 / nu;
  const cfl = opts.cfl || 0.4, dtFixed = opts.dt || 0;
  factorize(N);
  const kc = Math.floor(N / 3);
  const ic0 = opts.ic || 'tgv';
  const v0 = icVector(ic0, opts.icParams);
  // the GPU runner's fp32 flag (nslab_gpu.py:56-65): complex64 storage; machine epsilon of the working precision and
  // the spectral noise floor it implies (E is quadratic in the field)
  const fp32 = !!opts.fp32;
  if (fp32 && pool) throw new Error('FP32 solves are never pool slots');
  const eps = fp32 ? 1.1920929e-07 : 2.220446049250313e-16, specFloor = 100 * eps * eps;
  // a pool slot (Pool.admit) or a solve of its own
  let h, slot = -1;
  if (pool) [slot, h] = addon.poolAdmit(pool.h, N, IC_KIND[ic0], nu, cfl, dtFixed, v0[0], v0[1], v0[2], v0[3]);
  else h = addon.create(N, IC_KIND[ic0], nu, cfl, dtFixed, v0[0], v0[1], v0[2], v0[3], fp32 ? 1 : 0);
  const st = { t: 0, step: 0, dt: 0, N, Re, nu, kc, ic: ic0, icParams: opts.icParams || {}, version: VERSION,
    series: { t: [], E: [], Z: [], eps: [], omMax: [], uMax: [], dt: [], ebal: [], zbal: [], Pspec: [], Pal: [] },
    maxEbal: 0, maxZbal: 0, maxTnl: 0, divMax: 0, outputs: [] };
  let last = null;
  // device clock as last pushed, so assignments to st (checkpoint restore) reach the device before stepping
  let devClock = { t: 0, step: 0, E0: 0 };
  // host mirror of the state (allocated on first use): handed out by the _S getter, uploaded before the next device
  // operation
  let S = null;
  let sOut = false;
  const sp = new Spectral(N);

  const toDiag = a => ({ E: a[0], Z: a[1], P: a[2], eps: a[3], Tnl: a[4], Pspec: a[5], omMax: a[6], uMax: a[7] });
  function refreshLast() { last = toDiag(addon.last(h)); return last; }
  function sync() {
    if (sOut) { addon.setState(h, S, NS); sOut = false; refreshLast(); }
    if (st.t !== devClock.t || st.step !== devClock.step || st.E0 !== devClock.E0) {
      addon.setClock(h, st.t, st.step, st.E0);
      devClock = { t: st.t, step: st.step, E0: st.E0 };
    }
  }
  function record(dt) {
    const s = st.series; s.t.push(st.t); s.E.push(last.E); s.Z.push(last.Z); s.eps.push(last.eps); s.omMax.push(last.omMax); s.uMax.push(last.uMax); s.dt.push(dt); s.Pspec.push(last.Pspec); s.Pal.push(last.P);
  }
  function finishIC() {
    for (const k of Object.keys(st.series)) st.series[k].length = 0;
    st.outputs.length = 0; st.maxEbal = st.maxZbal = st.maxTnl = st.divMax = 0;
    st.t = 0; st.step = 0;
    refreshLast(); st.E0 = last.E; st.Z0 = last.Z;
    devClock = { t: 0, step: 0, E0: st.E0 };
    addon.setClock(h, 0, 0, st.E0);
    record(0);
    return last;
  }

  /** setIC (nslab.js:250-302) */
  function setIC(kind, prm) {
    prm = Object.assign({}, prm || {});
    const v = icVector(kind, prm);
    sOut = false;
    addon.setIC(h, IC_KIND[kind], v[0], v[1], v[2], v[3]);
    st.ic = kind; st.icParams = prm;
    return finishIC();
  }
  // Books `got` step records (21 values each: t dt step ebal zbal last[8] d0[8]) starting at r[base] as step()
  // leaves st (nslab.js:314-332).
  function book(r, base, got) {
    for (let i = 0; i < got; i++) {
      const o = base + i * 21;
      const d0 = toDiag(r.subarray(o + 13, o + 21));
      st.maxTnl = Math.max(st.maxTnl, Math.abs(d0.Tnl) / Math.max(d0.eps, 1e-300));
      st.t = r[o]; st.dt = r[o + 1]; st.step = r[o + 2];
      const ebal = r[o + 3], zbal = r[o + 4];
      st.maxEbal = Math.max(st.maxEbal, ebal); st.maxZbal = Math.max(st.maxZbal, zbal);
      last = toDiag(r.subarray(o + 5, o + 13));
      record(st.dt); st.series.ebal.push(ebal); st.series.zbal.push(zbal);
    }
    if (got > 0) {
      devClock = { t: st.t, step: st.step, E0: st.E0 };
      sOut = false;   // the device state moved on: the next _S read fetches it
    }
  }
  /** run(nSteps, tEnd) (nslab.js:334): steps on the device in chunks; the series are booked as record() does. */
  function run(nSteps, tEnd) {
    sync();
    let left = nSteps;
    while (left > 0) {
      const k = Math.min(left, 1024);
      const r = addon.run(h, k, tEnd != null ? 1 : 0, tEnd != null ? tEnd : 0);
      const got = r.length / 21;
      book(r, 0, got);
      left -= got;
      if (got < k) break;
    }
    return last;
  }
  function step() { return run(1); }
  function norms() { sync(); const a = addon.norms(h); return { E: a[0], Z: a[1], P: a[2] }; }
  function spectrum() { sync(); return Array.from(addon.spectrum(h)); }

  /** cutoffPileUp (nslab.js:557-564): the library's rule over a spectrum of this grid; null below the noise floor of
   *  the working precision (nslab_gpu.py:369-372 for an FP32 solve). */
  function cutoffPileUp(spec) { return addon.cutoffPileUp(kc, Float64Array.from(spec), fp32 ? eps : undefined); }

  /** diagnose (nslab.js:339-392): the whole computation runs on the GPU; the result object has the reference's keys. */
  function diagnose() {
    sync();
    const a = addon.diagnose(h, 1);
    refreshLast();
    const spec = Array.from(a.subarray(80));
    const hist = [0, 1, 2].map(e => Array.from(a.subarray(32 + 16 * e, 48 + 16 * e)));
    const out = { t: st.t, step: st.step, E: a[2], Z: a[3], P: a[4], eps: a[5], omMax: a[6], uMax: a[7], Pspec: a[8], Pphys: a[9], Tnl: a[10],
      skew: a[11], divMax: a[12], eta: a[13], kmaxEta: a[14], Rel: a[15], lambda: a[16], skewIso: a[17],
      align: [a[18], a[19], a[20]], alignHist: hist, spectrum: spec, uL3: a[21], pileUp: Number.isNaN(a[22]) ? null : a[22] };
    Object.assign(out, { zCentroid: a[23], zExtent: a[24], imageGap: a[25], zBands: a[26] });
    if (!Number.isNaN(a[27])) Object.assign(out, { omMaxI: a[27], omMaxIpos: [a[28], a[29], a[30]], omMaxNode: a[31] });
    st.divMax = Math.max(st.divMax, out.divMax);
    st.outputs.push(out);
    return out;
  }
  function interpMax(top = 24, iters = 12) {
    sync();
    const a = addon.interpMax(h, top, iters);
    return { omMaxI: a[0], omMaxIpos: [a[1], a[2], a[3]], omMaxNode: a[4] };
  }
  // physical fields of the current state (u, omega) and, after diagnose(), Q and the stretching density
  const FIELD = { u: 0, v: 1, w: 2, omx: 3, omy: 4, omz: 5, q: 6, stretch: 7 };
  function field(name) { sync(); const f = new Float64Array(N3); addon.field(h, FIELD[name], f); return f; }
  const fields = {};
  for (const k of Object.keys(FIELD)) Object.defineProperty(fields, k, { enumerable: true, get: () => field(k) });

  /** imageDiag (nslab.js:537-550): the enstrophy profile along z of the current vorticity, on the GPU. */
  function imageDiag() {
    sync();
    const a = addon.imageDiag(h);
    return { zCentroid: a[0], zExtent: a[1], imageGap: a[2], zBands: a[3] };
  }

  /** slice (nslab.js:566-583): N x N Float32 plane of a physical quantity, extracted on the GPU. */
  const SLICE_KIND = { vort: 0, speed: 1, ke: 2, q: 3, stretch: 4, u: 5, v: 6, w: 7, omx: 8, omy: 9, omz: 10 };
  function slice(kind, axis, index) {
    sync();
    const k = kind in SLICE_KIND ? SLICE_KIND[kind] : -1;
    return addon.slice(h, k, axis === 'z' ? 2 : axis === 'y' ? 1 : 0, index | 0);
  }
  function exactError() {
    sync();
    const a = addon.exactError(h);
    return a ? { linf: a[0], rel: a[1] } : null;
  }

  // what report.js reads of this solve
  const run_ = { N, kc, nu, Re, cfl, dtFixed, st, fp32, get last() { return last; }, spectrum, pileUp: cutoffPileUp };
  /** Health report (nslab.js:595-635). */
  function health(study) { return report.healthReport(run_, study); }
  /** Evidence dossier (nslab.js:637-645). */
  function dossier(extra) {
    return report.dossierOf(run_, 'Pocket Wind Tunnel · NSLab ' + VERSION + ' (MI355X ' + (fp32 ? 'FP32 storage' : 'FP64') + ', ' + addon.version() + ')', extra);
  }
  /** Frees the device memory now (a pool member is left to its pool); later calls on this solver throw. */
  function destroy() { if (!pool) addon.destroy(h); }

  finishIC();
  const solver = { st, get last() { return last; }, step, run, diagnose, spectrum, slice, exactError, health, dossier, setIC,
    norms, interpMax, imageDiag, fields, N, nu, Re, kc, _sp: sp, destroy, fp32, eps, specFloor,
    precision: fp32 ? 'float32' : 'float64',
    deviceBytes: () => addon.deviceBytes(h),
    profile: on => addon.profile(h, on ? 1 : 0), profileRead: () => addon.profileRead(h),
    _slot: slot, _book: book, _sync: sync };
  /** _S: the six state planes (nslab.js:184), copied from the device on access; writes to them reach the device
   *  before the next solver operation. */
  Object.defineProperty(solver, '_S', { enumerable: true, get() {
    if (!S) S = [0, 1, 2, 3, 4, 5].map(() => new Float64Array(NS));
    if (!sOut) { addon.getState(h, S, NS); sOut = true; }
    return S;
  } });
  return solver;
}

/** Refinement study verdict (nslab.js:652-666). */
const studySummary = report.studySummary;

// ------------------------------------------------------------------------------------------------ pools
/**
 * Solves of one N resident on the device and stepped together: one launch per pass covers every slot, one RK4
 * step of the pool is a captured hipGraph. admit() returns a solver object with the createSolver() surface; its
 * series are booked by collect(). A slot's results are bit-identical to the same solve run alone.
 */
class Pool {
  constructor(N, slots) {
    this.h = addon.poolCreate(N, slots);
    this.N = N; this.slots = slots;
    this.members = new Map();   // slot -> solver
  }
  static slotBytes(N) { return addon.poolSlotBytes(N); }
  info() { const a = addon.poolInfo(this.h); return { N: a[0], slots: a[1], used: a[2], graphed: a[3], bytes: a[4] }; }
  get free() { return this.slots - this.members.size; }
  /** createSolver(opts) into a free slot (throws "pool is full" when none is). */
  admit(opts) {
    const s = createSolver(Object.assign({}, opts, { N: this.N }), this);
    this.members.set(s._slot, s);
    return s;
  }
  /** The solver's next launches run up to n steps, stopping at tEnd (run(n, tEnd) semantics). */
  setTarget(s, n, tEnd) { s._sync(); addon.poolSetTarget(this.h, s._slot, n, tEnd != null ? 1 : 0, tEnd != null ? tEnd : 0); }
  /** Enqueue up to k (<= 64) steps of every slot with steps left; returns at once. */
  launch(k) {
    for (const s of this.members.values()) s._sync();
    this.k = k || 64;
    addon.poolLaunch(this.h, this.k);
  }
  /** True once the launched steps have finished (collect() will not wait). */
  ready() { return addon.poolReady(this.h); }
  /** Wait for the launched steps, book them into each member's series; returns the number of slots still running. */
  collect() {
    const k = this.k || 64, r = addon.poolCollect(this.h, k), head = 1 + this.slots;
    for (const [slot, s] of this.members) { const got = r[1 + slot]; if (got > 0) s._book(r, head + slot * k * 21, got); }
    return r[0];
  }
  release(s) { addon.poolRelease(this.h, s._slot); this.members.delete(s._slot); }
  /** Frees the pool's device memory now; its members must not be used afterwards. */
  destroy() { addon.poolDestroy(this.h); this.members.clear(); }
}

/**
 * Pools of different N stepped as one: one RK4 step of every pool is one captured hipGraph, so an ensemble step of
 * uneven grids is one graph launch. launch(k) steps every pool; then each pool's ready()/collect() as usual. Destroy
 * the group before any of its pools.
 */
class PoolGroup {
  constructor(pools) {
    this.pools = pools.slice();
    this.h = addon.groupCreate(this.pools.map(p => p.h));
  }
  launch(k) {
    for (const p of this.pools) { for (const s of p.members.values()) s._sync(); p.k = k || 64; }
    addon.groupLaunch(this.h, k || 64);
  }
  destroy() { addon.groupDestroy(this.h); }
}

function deviceMemory() { const a = addon.deviceMemory(); return { free: a[0], total: a[1] }; }

module.exports = { VERSION, IC_INFO, createSolver, studySummary, Spectral, fftPlan, fft1, factorize, Pool, PoolGroup, deviceMemory,
  addonVersion: addon.version() };
