// Known-answer validation of the MI355X NSLab host module (build-me-anything.github.io_amd/js/nslab.js), run under
// Node on a GPU box. The checks restate the reference's validation suite (pocket-wind-tunnel/test/validate-ns.js:12-125:
// FFT vs direct DFT, transform round trip, exact ABC and 2D Taylor-Green decay, RK4 order, Taylor-Green invariants and
// budgets, refinement ladder grading, solenoidality, interpolated maximum, cross-instrument tubes constants, image
// gap, L3 norm, pile-up, worst-instant health rows) and add parity checks against the golden outputs the reference
// wrote (tests/golden/*.json). Exit status 0 and "ALL PASS" on success.
'use strict';
const path = require('path');
const fs = require('fs');
const NS = require(path.join(__dirname, '..', '..', 'build-me-anything.github.io_amd', 'js', 'nslab.js'));
const golden = name => JSON.parse(fs.readFileSync(path.join(__dirname, '..', 'golden', name), 'utf8'));

let fails = 0;
const only = process.argv[2] || 'all';
function check(name, val, lo, hi) {
  const ok = val >= lo && val <= hi;
  if (!ok) fails++;
  const shown = typeof val === 'number' ? (Math.abs(val) < 1e-3 && val !== 0 ? val.toExponential(3) : val.toFixed(5)) : String(val);
  console.log(`${ok ? 'PASS' : 'FAIL'}  ${name.padEnd(70)} ${shown}   [${lo}, ${hi}]`);
}
const t0 = Date.now();
const want = g => only === 'all' || only === g;

if (want('fft')) {
  let worst = 0;
  for (const N of [12, 16, 24, 32, 48, 64, 96]) {
    const plan = NS.fftPlan(N);
    for (let i = 0; i < N; i++) { plan.xr[i] = Math.sin(1.3 * i) + 0.2 * i; plan.xi[i] = Math.cos(0.7 * i * i); }
    const xr = Float64Array.from(plan.xr), xi = Float64Array.from(plan.xi);
    const [Fr, Fi] = NS.fft1(plan, -1);
    for (let k = 0; k < N; k++) {
      let dr = 0, di = 0;
      for (let n = 0; n < N; n++) { const a = -2 * Math.PI * k * n / N; dr += xr[n] * Math.cos(a) - xi[n] * Math.sin(a); di += xr[n] * Math.sin(a) + xi[n] * Math.cos(a); }
      worst = Math.max(worst, Math.abs(Fr[k] - dr), Math.abs(Fi[k] - di));
    }
  }
  check('FFT: max |FFT - direct DFT| over N = 12..96', worst, 0, 1e-9);
  const N = 24, sp = new NS.Spectral(N), f = new Float64Array(N ** 3), g = new Float64Array(N ** 3), h = 2 * Math.PI / N;
  for (let l = 0; l < N; l++) for (let j = 0; j < N; j++) for (let i = 0; i < N; i++)
    f[(l * N + j) * N + i] = Math.sin(3 * i * h) * Math.cos(2 * j * h) + 0.5 * Math.cos(5 * l * h) * Math.sin(i * h) + 0.1;
  const re = new Float64Array(sp.NS), im = new Float64Array(sp.NS);
  sp.forward(f, re, im); sp.inverse(re, im, g);
  let e = 0; for (let i = 0; i < f.length; i++) e = Math.max(e, Math.abs(f[i] - g[i]));
  check('3D real transform round trip (24^3)', e, 0, 1e-12);
  check('3D transform: mean mode equals the field mean (0.1)', Math.abs(re[0] - 0.1), 0, 1e-14);
  let threw = 0; try { NS.createSolver({ N: 7 }); } catch (err) { threw = /2\^a·3\^b/.test(err.message) ? 1 : 0; }
  check('bad grid size throws the reference message', threw, 1, 1);
}

if (want('exact')) {
  const s = NS.createSolver({ N: 16, Re: 10, ic: 'abc', cfl: 0.4 }); s.run(1e6, 1.0);
  const ex = s.exactError(), d = s.diagnose();
  check('ABC exact solution: L-inf velocity error at t = 1', ex.linf, 0, 1e-10);
  check('ABC: energy matches 1.5 e^{-2 nu t}', Math.abs(d.E / (1.5 * Math.exp(-2 * 0.1 * s.st.t)) - 1), 0, 1e-10);
  check('ABC: divergence round-off', d.divMax, 0, 1e-12);
  check('ABC: nonlinear energy transfer vanishes (|T|/eps)', Math.abs(d.Tnl) / d.eps, 0, 1e-12);
  const gA = golden('exact.json').abc;
  check('ABC: step count to t = 1 equals the reference', s.st.step === gA.st.step ? 1 : 0, 1, 1);
  check('ABC: L-inf error equals the reference (abs)', Math.abs(ex.linf - gA.exact.linf), 0, 1e-13);
  const s2 = NS.createSolver({ N: 16, Re: 10, ic: 'tgv2d', cfl: 0.4 }); s2.run(1e6, 1.0);
  check('2D Taylor-Green exact decay: L-inf error at t = 1', s2.exactError().linf, 0, 1e-8);
  const e = [];
  for (const dt of [0.1, 0.05, 0.025]) { const q = NS.createSolver({ N: 16, Re: 10, ic: 'abc', dt }); q.run(1e6, 1.0); e.push(q.exactError().linf); }
  check('RK4 temporal order (dt -> dt/2)', Math.log2(e[0] / e[1]), 3.7, 4.3);
  check('RK4 temporal order (dt/2 -> dt/4)', Math.log2(e[1] / e[2]), 3.7, 4.3);
}

if (want('tgv')) {
  const s = NS.createSolver({ N: 32, Re: 1600, ic: 'tgv', cfl: 0.4 }); s.run(1e6, 2.0); const d = s.diagnose();
  check('TGV: initial energy 1/8', Math.abs(s.st.E0 - 0.125), 0, 1e-14);
  check('TGV: initial enstrophy 3/8', Math.abs(s.st.Z0 - 0.375), 0, 1e-14);
  check('TGV: RK4-consistent energy budget residual', s.st.maxEbal, 0, 1e-5);
  check('TGV: enstrophy budget residual', s.st.maxZbal, 0, 1e-4);
  check('TGV: stretching spectral vs physical (t = 2)', Math.abs(d.Pspec - d.Pphys) / Math.abs(d.Pspec), 0, 1e-6);
  check('TGV: direct skewness vanishes by symmetry', Math.abs(d.skew), 0, 0.02);
  check('TGV: initial enstrophy production is zero (Pspec at t=0)', Math.abs(s.st.series.Pspec[0]), 0, 1e-12);
  const g = golden('tgv32_cfl04_t2.json');
  check('TGV t=2: step count equals the reference', s.st.step === g.st.step ? 1 : 0, 1, 1);
  let worst = 0;
  for (const k of ['t', 'E', 'Z', 'eps', 'omMax', 'uMax', 'dt', 'Pal']) for (let i = 0; i < g.series[k].length; i++)
    worst = Math.max(worst, Math.abs(s.st.series[k][i] - g.series[k][i]) / Math.max(Math.abs(g.series[k][i]), 1e-300));
  check('TGV t=2: series equal the reference (max rel)', worst, 0, 1e-13);
  check('TGV t=2: Pphys equals the reference (rel)', Math.abs(d.Pphys - g.diag.Pphys) / Math.abs(g.diag.Pphys), 0, 1e-12);
  const peaks = []; let dEnd = null, sEnd = null;
  for (const N of [16, 24, 32]) {
    const q = NS.createSolver({ N, Re: 1600, ic: 'tgv', cfl: 0.4 }); q.run(1e6, 10.0);
    if (N === 32) { dEnd = q.diagnose(); sEnd = q; }
    const ser = q.st.series; let ip = 0; for (let i = 0; i < ser.eps.length; i++) if (ser.eps[i] > ser.eps[ip]) ip = i;
    peaks.push({ N, eps: ser.eps[ip], t: ser.t[ip], health: q.health().worst });
  }
  check('TGV ladder: eps_max increases with resolution', peaks[0].eps < peaks[1].eps && peaks[1].eps < peaks[2].eps ? 1 : 0, 1, 1);
  check('TGV ladder: under-resolved grids are graded FAIL', peaks.every(p => p.health === 'FAIL') ? 1 : 0, 1, 1);
  check('TGV 32^3: eps_max / Brachet 0.0126', peaks[2].eps / 0.0126, 0.5, 1.05);
  check('TGV 32^3: time of eps_max', peaks[2].t, 8, 10.5);
  check('TGV 32^3 t=10: vorticity aligns with e2', dEnd.align[1] > dEnd.align[0] && dEnd.align[1] > dEnd.align[2] ? 1 : 0, 1, 1);
  const ser = sEnd.st.series; let smin = 0;
  for (let i = 0; i < ser.t.length; i++) { const S = -(6 * Math.sqrt(15) / 7) * ser.Pspec[i] / Math.pow(2 * ser.Z[i], 1.5); if (S < smin) smin = S; }
  check('TGV 32^3: peak enstrophy-production skewness', smin, -0.8, -0.15);
  const lad = golden('ladder_re100.json'), levels = [];
  for (const L of lad.levels) {
    const q = NS.createSolver({ N: L.N, Re: 100, ic: 'tgv', cfl: 0.4 }); q.run(1e6, 3.0);
    const s2 = q.st.series; let ie = 0, io = 0;
    for (let k = 0; k < s2.eps.length; k++) { if (s2.eps[k] > s2.eps[ie]) ie = k; if (s2.omMax[k] > s2.omMax[io]) io = k; }
    levels.push({ label: `${L.N}³`, N: L.N, dt: 0, peak: { omMax: s2.omMax[io], t: s2.t[io] }, epsPeak: { v: s2.eps[ie], t: s2.t[ie] }, Eend: s2.E[s2.E.length - 1] });
  }
  const sum = NS.studySummary(levels, 'grid');
  check('studySummary verdict equals the reference', sum.verdict === lad.summary.verdict ? 1 : 0, 1, 1);
  check('studySummary text equals the reference', sum.text === lad.summary.text ? 1 : 0, 1, 1);
}

if (want('misc')) {
  const s = NS.createSolver({ N: 24, Re: 500, ic: 'tubes' }); s.run(20); const d = s.diagnose();
  check('vortex tubes: divergence-free evolved field', d.divMax, 0, 1e-10);
  check('vortex tubes: energy positive and finite', Number.isFinite(d.E) && d.E > 0 ? 1 : 0, 1, 1);
  const r = NS.createSolver({ N: 24, Re: 500, ic: 'random', icParams: { seed: 7, E0: 0.5 } });
  check('random field: energy scaled to E0 = 0.5', Math.abs(r.st.E0 - 0.5), 0, 1e-10);
  r.run(10); const dr = r.diagnose(); check('random field: divergence-free after stepping', dr.divMax, 0, 1e-10);
  // shifted Taylor-Green through the _S planes
  const N = 32, NH = N / 2 + 1, q = NS.createSolver({ N, Re: 1600, ic: 'tgv' });
  const h = 2 * Math.PI / N, d0 = [0.37 * h, 0.21 * h, 0.44 * h], kw = new Float64Array(N);
  for (let i = 0; i < N; i++) kw[i] = i <= N / 2 ? i : i - N;
  const S = q._S;
  for (let l = 0; l < N; l++) for (let j = 0; j < N; j++) for (let i = 0; i < NH; i++) {
    const idx = (l * N + j) * NH + i, ph = -(i * d0[0] + kw[j] * d0[1] + kw[l] * d0[2]), c = Math.cos(ph), sn = Math.sin(ph);
    for (let cc = 0; cc < 3; cc++) { const re0 = S[2 * cc][idx], im0 = S[2 * cc + 1][idx]; S[2 * cc][idx] = re0 * c - im0 * sn; S[2 * cc + 1][idx] = re0 * sn + im0 * c; }
  }
  const dq = q.diagnose();
  check('shifted TGV: grid maximum under-reports the exact 2', dq.omMax, 1.90, 1.999);
  check('shifted TGV: interpolated maximum recovers 2', Math.abs(dq.omMaxI - 2), 0, 1e-9);
  check('interpolated maximum >= grid maximum', dq.omMaxI >= dq.omMax ? 1 : 0, 1, 1);
  const t24 = NS.createSolver({ N: 24, Re: 4000, ic: 'tubes' }); const dt24 = t24.diagnose();
  check('tubes 24^3 interpolated max|omega| vs the GPU runner (7.73221)', Math.abs(dt24.omMaxI - 7.73221), 0, 2e-4);
  check('tubes 24^3 grid max|omega| vs the GPU runner (7.69712)', Math.abs(dt24.omMax - 7.69712), 0, 2e-4);
  const gt = golden('tubes.json').re4000.diag0;
  check('tubes 24^3 omMaxI equals the reference', Math.abs(dt24.omMaxI - gt.omMaxI), 0, 1e-9);
  check('tubes: periodic-image gap in z positive at t = 0', dt24.imageGap, 0.5, 6.0);
  check('tubes: image gap equals the reference', Math.abs(dt24.imageGap - gt.imageGap), 0, 1e-12);
  check('tubes: enstrophy band z-centroid at pi', Math.abs(dt24.zCentroid - Math.PI), 0, 1e-6);
  const tg = NS.createSolver({ N: 24, Re: 1600, ic: 'tgv' }); const dtg = tg.diagnose();
  check('Taylor-Green fills the box in z: image gap zero', dtg.imageGap, 0, 1e-12);
  const ref = (() => { const M = 96; let s3 = 0; for (let l = 0; l < M; l++) for (let j = 0; j < M; j++) for (let i = 0; i < M; i++) {
    const x = i * 2 * Math.PI / M, y = j * 2 * Math.PI / M, z = l * 2 * Math.PI / M;
    const u = Math.sin(x) * Math.cos(y) * Math.cos(z), v = -Math.cos(x) * Math.sin(y) * Math.cos(z);
    s3 += Math.pow(u * u + v * v, 1.5); } return Math.pow(s3 / (M * M * M), 1 / 3); })();
  check('TGV L3 norm against a direct quadrature', Math.abs(dtg.uL3 / ref - 1), 0, 5e-3);
  const sm = NS.createSolver({ N: 48, Re: 1600, ic: 'tgv' }); sm.run(20); const dsm = sm.diagnose();
  check('resolved TGV: no energy pile-up at the cutoff', dsm.pileUp == null ? 1 : dsm.pileUp, 0, 1.05);
  const hh = sm.health();
  check('health report exposes the worst-instant rows', hh.rows.filter(rw => /worst instant/.test(rw[0])).length, 3, 4);
  check('health worst verdict at least as severe as worstEnd', ({ PASS: 0, WARN: 1, FAIL: 2 }[hh.worst] >= { PASS: 0, WARN: 1, FAIL: 2 }[hh.worstEnd]) ? 1 : 0, 1, 1);
  const gh = golden('diag_misc.json').tgv48.health;
  check('TGV 48^3 health grades equal the reference', hh.rows.every((rw, i) => rw[2] === gh.rows[i][2]) && hh.rows.length === gh.rows.length ? 1 : 0, 1, 1);
}

if (want('config1')) {
  // The CPU-runnable configuration: TGV 32^3, nu = 1/1600, fixed dt = 0.01, 200 steps; fields, series and health
  const g = golden('tgv32_dt01.json');
  const s = NS.createSolver({ N: 32, Re: 1600, ic: 'tgv', dt: 0.01 });
  s.run(200);
  let worst = 0;
  for (const k of ['t', 'E', 'Z', 'eps', 'omMax', 'uMax', 'dt', 'Pal', 'Pspec']) for (let i = 0; i < g.series[k].length; i++)
    worst = Math.max(worst, Math.abs(s.st.series[k][i] - g.series[k][i]) / Math.max(...g.series[k].map(Math.abs)));
  check('config 1: series equal the reference after 200 steps (max rel)', worst, 0, 1e-13);
  const hn = s.health(), gn = g.healthNoDiag;
  check('config 1: health verdict (before diagnose) equals the reference', hn.worst === gn.worst ? 1 : 0, 1, 1);
  const d = s.diagnose(), hd = s.health(), gd = g.health;
  check('config 1: health verdict equals the reference', hd.worst === gd.worst && hd.worstEnd === gd.worstEnd ? 1 : 0, 1, 1);
  check('config 1: every health grade equals the reference', hd.rows.every((rw, i) => rw[2] === gd.rows[i][2]) && hd.rows.length === gd.rows.length ? 1 : 0, 1, 1);
  // rows whose value is physics must print identically (the divergence, |T|/eps, budget residuals and the
  // spectral-vs-physical stretching difference are roundoff-sized here: 1e-16..1e-14, graded above)
  const physics = ['grid', 'Δt', 'CFL (u_max Δt/Δx)', 'resolution kmax·η', 'spectral tail E(kmax)/E(peak)',
    'cutoff pile-up E(k)/E(0.8kmax)', 'max |ω|', 'enstrophy Z', 'stretching ⟨ω·S·ω⟩', 'dissipation ε = 2νZ', '‖u‖_L³ (ESS criterion)', 'periodic-image gap in z'];
  const same = physics.every(name => {
    const a = hd.rows.find(rw => rw[0] === name), b = gd.rows.find(rw => rw[0] === name);
    if (!(a && b && a[1] === b[1])) console.log('   row differs:', name, a && a[1], '| reference:', b && b[1]);
    return a && b && a[1] === b[1];
  });
  check('config 1: physics rows of the health report print identically', same ? 1 : 0, 1, 1);
  check('config 1: budget rows PASS', ['energy budget residual (RK4-consistent)', 'enstrophy budget residual', 'divergence L∞', 'nonlinear energy transfer |T|/ε']
    .every(name => hd.rows.find(rw => rw[0] === name)[2] === 'PASS') ? 1 : 0, 1, 1);
  check('config 1: diagnose E equals the reference (rel)', Math.abs(d.E - g.diag.E) / g.diag.E, 0, 1e-13);
  const dos = s.dossier({});
  check('dossier carries parameters, series, health', dos.parameters.N === 32 && dos.series.E.length === 201 && dos.health.rows.length > 10 ? 1 : 0, 1, 1);
  // checkpoint round trip through _S and st, as a batch runner restores one
  const planes = s._S.map(a => Float64Array.from(a));
  const meta = JSON.parse(JSON.stringify({ t: s.st.t, step: s.st.step, series: s.st.series, maxEbal: s.st.maxEbal, maxZbal: s.st.maxZbal, maxTnl: s.st.maxTnl, divMax: s.st.divMax, E0: s.st.E0, Z0: s.st.Z0 }));
  s.run(5); const Ea = s.last.E;
  const r = NS.createSolver({ N: 32, Re: 1600, ic: 'tgv', dt: 0.01 });
  for (let c = 0; c < 6; c++) r._S[c].set(planes[c]);
  Object.assign(r.st, meta);
  r.run(5);
  check('checkpoint restore continues bit-identically', r.last.E === Ea && r.st.step === s.st.step && r.st.t === s.st.t ? 1 : 0, 1, 1);
}

console.log(`\n${fails === 0 ? 'ALL PASS' : fails + ' FAILURES'}  (${((Date.now() - t0) / 1000).toFixed(1)} s)`);
process.exit(fails ? 1 : 0);
