/*
 * NSLab on MI355X — Node host module. A drop-in for pocket-wind-tunnel/src/nslab.js: the same exports
 * ({ VERSION, IC_INFO, createSolver, studySummary, Spectral, fftPlan, fft1, factorize }, nslab.js:668) and the same
 * solver object (nslab.js:648), with every transform, right-hand side, RK4 step and diagnostic computed in FP64 on the
 * GPU by libnslab_hip.so through the N-API addon. This file keeps the host-side bookkeeping the reference keeps on
 * the host: the mutable `st` object and its series (nslab.js:188-190, 300-302), the health report (nslab.js:595-635),
 * the dossier (nslab.js:637-645) and the refinement verdict (nslab.js:652-666).
 *
 * There is no CPU path: if the addon cannot be loaded the require() throws.
 */
'use strict';
const path = require('path');
const addon = require(path.join(__dirname, '..', 'lib', 'nslab_napi.node'));

const VERSION = '0.1.2';
const TWO_PI = 2 * Math.PI;
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
  // a pool slot (Pool.admit) or a solve of its own
  let h, slot = -1;
  if (pool) [slot, h] = addon.poolAdmit(pool.h, N, IC_KIND[ic0], nu, cfl, dtFixed, v0[0], v0[1], v0[2], v0[3]);
  else h = addon.create(N, IC_KIND[ic0], nu, cfl, dtFixed, v0[0], v0[1], v0[2], v0[3]);
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
    if (got > 0) devClock = { t: st.t, step: st.step, E0: st.E0 };
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

  const SPEC_FLOOR = 100 * Number.EPSILON * Number.EPSILON;
  function cutoffPileUp(spec) {
    const k8 = Math.round(0.8 * kc); let peak = 0; for (let k = 1; k < spec.length; k++) peak = Math.max(peak, spec[k]);
    if (!(spec[k8] > SPEC_FLOOR * peak)) return null;
    let m = 0; for (let k = k8; k <= kc; k++) m = Math.max(m, spec[k] / spec[k8]);
    return m;
  }

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

  /** imageDiag (nslab.js:537-550), from the vorticity field. */
  function imageDiag() {
    const ox = field('omx'), oy = field('omy'), oz = field('omz'), Ez = new Float64Array(N);
    for (let l = 0; l < N; l++) { let s = 0; for (let n = l * N * N; n < (l + 1) * N * N; n++) s += ox[n] * ox[n] + oy[n] * oy[n] + oz[n] * oz[n]; Ez[l] = s; }
    let mx = 0, cs = 0, sn = 0; for (let l = 0; l < N; l++) { mx = Math.max(mx, Ez[l]); const z = l * TWO_PI / N; cs += Ez[l] * Math.cos(z); sn += Ez[l] * Math.sin(z); }
    const zc = ((Math.atan2(sn, cs) % TWO_PI) + TWO_PI) % TWO_PI;
    const on = Array.from(Ez, v => v > 0.01 * mx); const runs = []; let start = null;
    for (let l = 0; l < N; l++) { if (on[l] && start === null) start = l; if (!on[l] && start !== null) { runs.push([start, l]); start = null; } }
    if (start !== null) runs.push([start, N]);
    let rr = runs; if (runs.length > 1 && on[0] && on[N - 1]) rr = [[runs[runs.length - 1][0], runs[0][1] + N]].concat(runs.slice(1, -1));
    const ext = Math.min((rr.length ? Math.max(...rr.map(r => r[1] - r[0])) : N) * TWO_PI / N, TWO_PI);
    return { zCentroid: zc, zExtent: ext, imageGap: Math.max(TWO_PI - ext, 0), zBands: rr.length };
  }

  /** slice (nslab.js:566-583): N x N Float32 slice of a physical field. */
  function slice(kind, axis, index) {
    const out = new Float32Array(N * N);
    let val;
    if (kind === 'vort' || kind === 'speed' || kind === 'ke') {
      const a = kind === 'vort' ? [field('omx'), field('omy'), field('omz')] : [field('u'), field('v'), field('w')];
      val = kind === 'vort' ? n => Math.sqrt(a[0][n] * a[0][n] + a[1][n] * a[1][n] + a[2][n] * a[2][n])
        : kind === 'speed' ? n => Math.sqrt(a[0][n] * a[0][n] + a[1][n] * a[1][n] + a[2][n] * a[2][n])
          : n => 0.5 * (a[0][n] * a[0][n] + a[1][n] * a[1][n] + a[2][n] * a[2][n]);
    } else if (kind in FIELD) { const f = field(kind); val = n => f[n]; }
    else val = () => 0;
    const ix = Math.max(0, Math.min(N - 1, index | 0));
    for (let b = 0; b < N; b++) for (let a = 0; a < N; a++) {
      const n = axis === 'z' ? (ix * N + b) * N + a : axis === 'y' ? (b * N + ix) * N + a : (b * N + a) * N + ix;
      out[b * N + a] = val(n);
    }
    return out;
  }
  function exactError() {
    sync();
    const a = addon.exactError(h);
    return a ? { linf: a[0], rel: a[1] } : null;
  }

  /** Health report (nslab.js:595-635): the reference's rows, thresholds and formats. */
  function health(study) {
    const d = last, o = st.outputs.length ? st.outputs[st.outputs.length - 1] : null;
    const eta = Math.pow(nu * nu * nu / Math.max(d.eps, 1e-300), 0.25), kmax = Math.sqrt(3) * kc, ke = kmax * eta;
    const spec = o ? o.spectrum : spectrum(); let kp = 0; for (let k = 1; k < spec.length; k++) if (spec[k] > spec[kp]) kp = k;
    const tail = spec[kc] / Math.max(spec[kp], 1e-300), pile = cutoffPileUp(spec);
    const cflNow = d.uMax * st.dt / (TWO_PI / N);
    const grade = (v, good, warn, lowerIsBetter = true) => (lowerIsBetter ? (v <= good ? 'PASS' : v <= warn ? 'WARN' : 'FAIL') : (v >= good ? 'PASS' : v >= warn ? 'WARN' : 'FAIL'));
    const sdOf = q => Math.abs(q.Pspec - q.Pphys) / Math.max(Math.abs(q.Pspec), 2 * nu * q.P, 1e-300);
    const rows = [
      ['grid', `${N}³ (kmax ${kc}, dealiased 2/3)`, ''],
      ['Δt', st.dt.toExponential(3) + (dtFixed ? ' fixed' : ' adaptive'), ''],
      ['CFL (u_max Δt/Δx)', cflNow.toFixed(3), grade(cflNow, 0.8, 1.2)],
      ['divergence L∞', st.divMax.toExponential(2), grade(st.divMax, 1e-10, 1e-6)],
      ['nonlinear energy transfer |T|/ε', st.maxTnl.toExponential(2), grade(st.maxTnl, 1e-9, 1e-6)],
      ['energy budget residual (RK4-consistent)', st.maxEbal.toExponential(2), grade(st.maxEbal, 1e-5, 1e-3)],
      ['enstrophy budget residual', st.maxZbal.toExponential(2), grade(st.maxZbal, 1e-4, 1e-2)],
      ['resolution kmax·η', ke.toFixed(2), grade(ke, 1.0, 0.5, false)],
      ['spectral tail E(kmax)/E(peak)', tail.toExponential(2), grade(tail, 1e-4, 1e-2)],
      ['stretching: spectral vs physical', o ? sdOf(o).toExponential(2) : '—', o ? grade(sdOf(o), 1e-2, 1e-1) : ''],
      ['cutoff pile-up E(k)/E(0.8kmax)', pile != null ? pile.toFixed(2) : '—', pile != null ? grade(pile, 1.2, 2.0) : ''],
      ['max |ω|', d.omMax.toFixed(4) + (o && o.omMaxI != null ? ` (interpolated ${o.omMaxI.toFixed(4)})` : ''), ''], ['enstrophy Z', d.Z.toFixed(5), ''], ['stretching ⟨ω·S·ω⟩', d.Pspec.toFixed(5), ''], ['dissipation ε = 2νZ', d.eps.toExponential(4), ''],
      ['‖u‖_L³ (ESS criterion)', o && o.uL3 != null ? o.uL3.toFixed(5) : '—', ''],
      ['periodic-image gap in z', o && o.imageGap != null ? o.imageGap.toFixed(2) : '—', ''],
      ['grid convergence', study && study.grid ? study.grid.text : 'not run', study && study.grid ? study.grid.verdict : 'N/A'],
      ['time-step convergence', study && study.time ? study.time.text : 'not run', study && study.time ? study.time.verdict : 'N/A'],
    ];
    const verdict = rs => (rs.some(r => r[2] === 'FAIL') ? 'FAIL' : rs.some(r => r[2] === 'WARN') ? 'WARN' : 'PASS');
    const worstEnd = verdict(rows);
    const snaps = st.outputs;
    if (snaps.length) {
      const tailOf = q => { let kq = 0; for (let k = 1; k < q.spectrum.length; k++) if (q.spectrum[k] > q.spectrum[kq]) kq = k; return q.spectrum[kc] / Math.max(q.spectrum[kq], 1e-300); };
      const wKe = snaps.reduce((a, b) => (b.kmaxEta < a.kmaxEta ? b : a)), wT = snaps.reduce((a, b) => (tailOf(b) > tailOf(a) ? b : a)), wS = snaps.reduce((a, b) => (sdOf(b) > sdOf(a) ? b : a));
      const wP = snaps.filter(q => q.pileUp != null).reduce((a, b) => ((a == null || b.pileUp > a.pileUp) ? b : a), null);
      rows.push(['worst instant: kmax·η', `${wKe.kmaxEta.toFixed(2)} at t ${wKe.t.toFixed(2)}`, grade(wKe.kmaxEta, 1.0, 0.5, false)]);
      rows.push(['worst instant: spectral tail', `${tailOf(wT).toExponential(2)} at t ${wT.t.toFixed(2)}`, grade(tailOf(wT), 1e-4, 1e-2)]);
      rows.push(['worst instant: stretching spectral vs physical', `${sdOf(wS).toExponential(2)} at t ${wS.t.toFixed(2)}`, grade(sdOf(wS), 1e-2, 1e-1)]);
      if (wP) rows.push(['worst instant: cutoff pile-up', `${wP.pileUp.toFixed(2)} at t ${wP.t.toFixed(2)}`, grade(wP.pileUp, 1.2, 2.0)]);
    }
    return { rows, worst: verdict(rows), worstEnd, note: 'A run that does not PASS resolution, budgets and refinement cannot support any claim about the equations. Numerical growth ≠ mathematical singularity. The verdict includes the worst archived snapshot; worstEnd is the end-of-run verdict alone.' };
  }
  /** Evidence dossier (nslab.js:637-645). */
  function dossier(extra) {
    return Object.assign({
      instrument: 'Pocket Wind Tunnel · NSLab ' + VERSION + ' (MI355X FP64, ' + addon.version() + ')', build: null,
      equations: '3D incompressible Navier–Stokes, periodic [0,2π]³, Re = 1/ν, Fourier pseudo-spectral (2/3 dealiased, rotational form), RK4',
      parameters: { N, Re, nu, cfl, dt: dtFixed || 'adaptive', ic: st.ic, icParams: st.icParams, kc },
      state: { t: st.t, steps: st.step }, series: st.series, outputs: st.outputs, health: health(extra && extra.study),
      disclaimer: 'Numerical evidence only. Nothing here is a proof of global regularity or of finite-time blow-up of the Navier–Stokes equations; every feature must survive spatial and temporal refinement before it is even a conjecture.',
    }, extra || {});
  }

  finishIC();
  const solver = { st, get last() { return last; }, step, run, diagnose, spectrum, slice, exactError, health, dossier, setIC,
    norms, interpMax, imageDiag, fields, N, nu, Re, kc, _sp: sp,
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
function studySummary(levels, key) {
  if (levels.length < 2) return null;
  const rel = (a, b) => Math.abs(a - b) / Math.max(Math.abs(b), 1e-300);
  const rows = []; let lastChange = null, prevChange = null;
  for (let i = 1; i < levels.length; i++) {
    const a = levels[i - 1], b = levels[i];
    const c = { from: a.label, to: b.label, omMax: rel(a.peak.omMax, b.peak.omMax), eps: rel(a.epsPeak.v, b.epsPeak.v), Eend: rel(a.Eend, b.Eend), tPeak: Math.abs(a.epsPeak.t - b.epsPeak.t) };
    rows.push(c); prevChange = lastChange; lastChange = Math.max(c.omMax, c.eps, c.Eend);
  }
  const order = prevChange != null && lastChange > 0 ? Math.log(prevChange / lastChange) / Math.log(key === 'time' ? 2 : levels[levels.length - 1].N / levels[levels.length - 2].N) : null;
  const verdict = lastChange < 0.01 && (prevChange == null || lastChange < prevChange) ? 'PASS' : lastChange < 0.05 ? 'WARN' : 'FAIL';
  const last = rows[rows.length - 1];
  const text = `${levels.map(l => l.label).join(' → ')}: last change ω_max ${(last.omMax * 100).toFixed(2)} %, ε_peak ${(last.eps * 100).toFixed(2)} %, E(t_end) ${(last.Eend * 100).toFixed(3)} %` + (order != null ? `, observed order ${order.toFixed(2)}` : '');
  return { rows, verdict, text, order };
}

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
  launch(k) { this.k = k || 64; addon.poolLaunch(this.h, this.k); }
  /** True once the launched steps have finished (collect() will not wait). */
  ready() { return addon.poolReady(this.h); }
  /** Wait for the launched steps, book them into each member's series; returns the number of slots still running. */
  collect() {
    const k = this.k || 64, r = addon.poolCollect(this.h, k), head = 1 + this.slots;
    for (const [slot, s] of this.members) { const got = r[1 + slot]; if (got > 0) s._book(r, head + slot * k * 21, got); }
    return r[0];
  }
  release(s) { addon.poolRelease(this.h, s._slot); this.members.delete(s._slot); }
}

function deviceMemory() { const a = addon.deviceMemory(); return { free: a[0], total: a[1] }; }

module.exports = { VERSION, IC_INFO, createSolver, studySummary, Spectral, fftPlan, fft1, factorize, Pool, deviceMemory,
  addonVersion: addon.version() };
