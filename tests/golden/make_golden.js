// Writes the golden fixtures under tests/golden/ by running the NSLab reference solver under Node.
//
// Usage (only where the reference checkout is present; the fixtures it writes are committed):
//   node tests/golden/make_golden.js /root/reference/pocket-wind-tunnel/src/nslab.js [small|long|seeded [sizes]]
//
// What is written, and why no closed form or archived file pins it instead:
//   * full spectral states after 0/100/200 fixed-dt RK4 steps (TGV 32^3, nu = 1/1600, dt = 0.01): the archive holds
//     no field dumps, and the fixed-dt series are not archived (all archived runs are CFL-adaptive);
//   * random-IC realisations (Mulberry32 + Box-Muller + Hermitian fill, loop-order dependent);
//   * diagnose()/health() output for the same cases;
//   * 'long': the 64^3 TGV dissipation series at dt = 0.004 to t = 10 (2500 steps).
// States are stored compact: only the modes inside the 2/3 mask (|kx|,|ky|,|kz| <= floor(N/3)), in the loop order
// kz-index, ky-index, kx, as 6 little-endian float64 planes (ux re, ux im, uy re, uy im, uz re, uz im), gzip'd.
'use strict';
const fs = require('fs'), path = require('path'), zlib = require('zlib');
const refPath = process.argv[2];
const mode = process.argv[3] || 'small';
if (!refPath) { console.error('usage: make_golden.js <path to nslab.js> [small|long]'); process.exit(2); }
const NS = require(path.resolve(refPath));
const OUT = __dirname;

function compactState(N, S) {
  const NH = N / 2 + 1, kc = Math.floor(N / 3);
  const kw = j => (j <= N / 2 ? j : j - N);
  const idx = [];
  for (let l = 0; l < N; l++) for (let j = 0; j < N; j++) for (let i = 0; i < NH; i++)
    if (Math.abs(kw(l)) <= kc && Math.abs(kw(j)) <= kc && i <= kc) idx.push((l * N + j) * NH + i);
  const out = new Float64Array(6 * idx.length);
  for (let c = 0; c < 6; c++) for (let m = 0; m < idx.length; m++) out[c * idx.length + m] = S[c][idx[m]];
  return out;
}
function writeF64(name, arr) {
  const buf = Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength);
  fs.writeFileSync(path.join(OUT, name), zlib.gzipSync(buf, { level: 9 }));
}
function writeJSON(name, obj) { fs.writeFileSync(path.join(OUT, name), JSON.stringify(obj)); }
const stScalars = s => ({ t: s.st.t, step: s.st.step, dt: s.st.dt, E0: s.st.E0, Z0: s.st.Z0, maxEbal: s.st.maxEbal,
  maxZbal: s.st.maxZbal, maxTnl: s.st.maxTnl, divMax: s.st.divMax });
const stripDiag = d => { const o = Object.assign({}, d); return o; };

if (mode === 'small') {
  // 1. 1D FFT against the reference's own butterflies, the input of its DFT check
  const fft = {};
  for (const N of [12, 16, 24, 32, 48, 64, 96]) {
    const plan = NS.fftPlan(N), xr = [], xi = [];
    for (let i = 0; i < N; i++) { xr.push(Math.sin(1.3 * i) + 0.2 * i); xi.push(Math.cos(0.7 * i * i)); }
    plan.xr.set(xr); plan.xi.set(xi); const [Fr, Fi] = NS.fft1(plan, -1);
    const fwd = [Array.from(Fr), Array.from(Fi)];
    plan.xr.set(xr); plan.xi.set(xi); const [Br, Bi] = NS.fft1(plan, 1);
    fft[N] = { factors: NS.factorize(N), xr, xi, fwd, inv: [Array.from(Br), Array.from(Bi)] };
  }
  writeJSON('fft1.json', fft);

  // 2. 3D forward transform of the reference's 24^3 test field (full half-spectrum layout)
  {
    const N = 24, sp = new NS.Spectral(N), f = new Float64Array(N ** 3), h = 2 * Math.PI / N;
    for (let l = 0; l < N; l++) for (let j = 0; j < N; j++) for (let i = 0; i < N; i++)
      f[(l * N + j) * N + i] = Math.sin(3 * i * h) * Math.cos(2 * j * h) + 0.5 * Math.cos(5 * l * h) * Math.sin(i * h) + 0.1;
    const re = new Float64Array(sp.NS), im = new Float64Array(sp.NS); sp.forward(f, re, im);
    const both = new Float64Array(2 * sp.NS); both.set(re, 0); both.set(im, sp.NS);
    writeF64('spectral24_fwd.f64.gz', both);
  }

  // 3. Taylor-Green 32^3, nu = 1/1600, fixed dt = 0.01, 200 steps (the CPU-runnable configuration)
  {
    const s = NS.createSolver({ N: 32, Re: 1600, ic: 'tgv', dt: 0.01 });
    const t0 = Date.now();
    writeF64('tgv32_dt01_s000.f64.gz', compactState(32, s._S));
    const last0 = stripDiag(s.last);
    s.run(100); writeF64('tgv32_dt01_s100.f64.gz', compactState(32, s._S));
    const last100 = stripDiag(s.last), st100 = stScalars(s);
    s.run(100); writeF64('tgv32_dt01_s200.f64.gz', compactState(32, s._S));
    const last200 = stripDiag(s.last), st200 = stScalars(s);
    const healthNoDiag = s.health();
    const diag = s.diagnose();
    const health = s.health();
    writeJSON('tgv32_dt01.json', { opts: { N: 32, Re: 1600, ic: 'tgv', dt: 0.01 }, last0, last100, st100, last200, st200,
      series: s.st.series, healthNoDiag, diag, health, secs: (Date.now() - t0) / 1000 });
    console.log('tgv32 dt 0.01 x 200:', (Date.now() - t0) / 1000, 's');
  }

  // 4. random solenoidal ICs at 24^3 (E0 = 0.5) for three seeds, and 10 adaptive steps of seed 7 at Re 500
  {
    const out = {};
    for (const seed of [1, 7, 42]) {
      const r = NS.createSolver({ N: 24, Re: 500, ic: 'random', icParams: { seed, E0: 0.5 } });
      writeF64(`random24_seed${seed}_s000.f64.gz`, compactState(24, r._S));
      out[seed] = { last0: stripDiag(r.last), E0: r.st.E0, Z0: r.st.Z0 };
      if (seed === 7) {
        r.run(10); writeF64('random24_seed7_s010.f64.gz', compactState(24, r._S));
        out[seed].last10 = stripDiag(r.last); out[seed].st10 = stScalars(r); out[seed].series = r.st.series;
        out[seed].diag10 = r.diagnose();
      }
    }
    // a random field with a non-default k0 and a non-default energy
    const q = NS.createSolver({ N: 32, Re: 2000, ic: 'random', icParams: { seed: 3, k0: 6, E0: 0.25 } });
    writeF64('random32_seed3_k06_s000.f64.gz', compactState(32, q._S));
    out['32_seed3_k06'] = { last0: stripDiag(q.last), E0: q.st.E0, Z0: q.st.Z0 };
    writeJSON('random.json', out);
  }

  // 5. vortex tubes: 24^3 Re 4000 at t = 0 (state + diagnose), 24^3 Re 500 for 20 adaptive steps
  {
    const t = NS.createSolver({ N: 24, Re: 4000, ic: 'tubes' });
    writeF64('tubes24_s000.f64.gz', compactState(24, t._S));
    const d0 = t.diagnose();
    const u = NS.createSolver({ N: 24, Re: 500, ic: 'tubes' }); u.run(20);
    writeF64('tubes24_re500_s020.f64.gz', compactState(24, u._S));
    const du = u.diagnose();
    const p = NS.createSolver({ N: 32, Re: 2000, ic: 'tubes', icParams: { amp: 5, sigma: 0.3, sep: 0.8, pert: 0.1 } });
    writeF64('tubes32_params_s000.f64.gz', compactState(32, p._S));
    writeJSON('tubes.json', { re4000: { last0: t.last, E0: t.st.E0, Z0: t.st.Z0, diag0: d0 },
      re500: { series: u.st.series, st: stScalars(u), last: u.last, diag: du, health: u.health() },
      params32: { icParams: { amp: 5, sigma: 0.3, sep: 0.8, pert: 0.1 }, last0: p.last, E0: p.st.E0, Z0: p.st.Z0 } });
  }

  // 6. exact solutions: ABC and 2D Taylor-Green, N 16, Re 10, to t = 1 (adaptive CFL 0.4); ABC at fixed dt ladder
  {
    const a = NS.createSolver({ N: 16, Re: 10, ic: 'abc', cfl: 0.4 }); a.run(1e6, 1.0);
    const aex = a.exactError(); const ad = a.diagnose();
    const b = NS.createSolver({ N: 16, Re: 10, ic: 'tgv2d', cfl: 0.4 }); b.run(1e6, 1.0);
    const bex = b.exactError();
    const ladder = [];
    for (const dt of [0.1, 0.05, 0.025]) { const q = NS.createSolver({ N: 16, Re: 10, ic: 'abc', dt }); q.run(1e6, 1.0); ladder.push({ dt, steps: q.st.step, t: q.st.t, linf: q.exactError().linf }); }
    const c = NS.createSolver({ N: 16, Re: 10, ic: 'abc', icParams: { A: 1, B: 0.7, C: 0.3 }, cfl: 0.4 }); c.run(1e6, 0.5);
    writeJSON('exact.json', { abc: { series: a.st.series, st: stScalars(a), exact: aex, diag: ad, health: a.health() },
      tgv2d: { series: b.st.series, st: stScalars(b), exact: bex },
      abcLadder: ladder, abcParams: { icParams: { A: 1, B: 0.7, C: 0.3 }, series: c.st.series, st: stScalars(c), exact: c.exactError() } });
  }

  // 7. Taylor-Green 32^3 Re 1600, CFL 0.4 to t = 2 (budgets, stretching two ways)
  {
    const s = NS.createSolver({ N: 32, Re: 1600, ic: 'tgv', cfl: 0.4 }); s.run(1e6, 2.0); const d = s.diagnose();
    writeF64('tgv32_cfl04_t2.f64.gz', compactState(32, s._S));
    writeJSON('tgv32_cfl04_t2.json', { series: s.st.series, st: stScalars(s), diag: d, health: s.health() });
  }

  // 8. shifted Taylor-Green (sub-cell phase shift of the state) for the interpolated maximum; 48^3 pile-up
  {
    const N = 32, NH = N / 2 + 1, s = NS.createSolver({ N, Re: 1600, ic: 'tgv' });
    const h = 2 * Math.PI / N, d0 = [0.37 * h, 0.21 * h, 0.44 * h], kw = new Float64Array(N);
    for (let i = 0; i < N; i++) kw[i] = i <= N / 2 ? i : i - N;
    const S = s._S;
    for (let l = 0; l < N; l++) for (let j = 0; j < N; j++) for (let i = 0; i < NH; i++) {
      const idx = (l * N + j) * NH + i, ph = -(i * d0[0] + kw[j] * d0[1] + kw[l] * d0[2]), c = Math.cos(ph), sn = Math.sin(ph);
      for (let cc = 0; cc < 3; cc++) { const re0 = S[2 * cc][idx], im0 = S[2 * cc + 1][idx]; S[2 * cc][idx] = re0 * c - im0 * sn; S[2 * cc + 1][idx] = re0 * sn + im0 * c; }
    }
    writeF64('tgv32_shifted.f64.gz', compactState(32, S));
    const d = s.diagnose();
    const sm = NS.createSolver({ N: 48, Re: 1600, ic: 'tgv' }); sm.run(20); const dsm = sm.diagnose();
    writeJSON('diag_misc.json', { shifted: d, tgv48: { diag: dsm, health: sm.health(), st: stScalars(sm) } });
  }

  // 9. a 3-level TGV grid ladder at Re 100 (16/24/32 to t = 3) and its studySummary verdict
  {
    const levels = [];
    for (const N of [16, 24, 32]) {
      const q = NS.createSolver({ N, Re: 100, ic: 'tgv', cfl: 0.4 }); q.run(1e6, 3.0);
      const s = q.st.series; let ie = 0, io = 0;
      for (let k = 0; k < s.eps.length; k++) { if (s.eps[k] > s.eps[ie]) ie = k; if (s.omMax[k] > s.omMax[io]) io = k; }
      levels.push({ label: `${N}³`, N, dt: 0, peak: { omMax: s.omMax[io], t: s.t[io] }, epsPeak: { v: s.eps[ie], t: s.t[ie] }, Eend: s.E[s.E.length - 1], steps: q.st.step });
    }
    writeJSON('ladder_re100.json', { levels, summary: NS.studySummary(levels, 'grid') });
  }
} else if (mode === 'seeded') {
  // Seeded random solenoidal fields (k0 4, E0 0.5, Re 1000, CFL 0.5) stepped 100 times at every grid the ensemble and
  // the ladders use below 192^3: the size-specialised kernels of 32..128 against the reference on a field with no
  // symmetry. 32^3 and 48^3 keep the whole state; 64..128 keep every `stride`-th compact mode (about 2000) and the
  // series.
  const sizes = (process.argv[4] || '32,48,64,96,128').split(',').map(Number);
  for (const N of sizes) {
    const seed = 100 + N, t0 = Date.now();
    const s = NS.createSolver({ N, Re: 1000, cfl: 0.5, ic: 'random', icParams: { seed, k0: 4, E0: 0.5 } });
    const c0 = compactState(N, s._S);
    s.run(100);
    const c1 = compactState(N, s._S);
    const nm = c1.length / 6, stride = N <= 48 ? 1 : Math.floor(nm / 2000);
    const pick = c => { const o = new Float64Array(6 * Math.ceil(nm / stride)); let q = 0;
      for (let p = 0; p < 6; p++) for (let m = 0; m < nm; m += stride) o[q++] = c[p * nm + m]; return o; };
    writeF64(`random${N}_s${seed}_s000.f64.gz`, pick(c0));
    writeF64(`random${N}_s${seed}_s100.f64.gz`, pick(c1));
    writeJSON(`random${N}_s${seed}.json`, { opts: { N, Re: 1000, cfl: 0.5, ic: 'random', icParams: { seed, k0: 4, E0: 0.5 } },
      stride, modes: nm, st: stScalars(s), last: s.last, series: s.st.series, secs: (Date.now() - t0) / 1000 });
    console.log(`random ${N}^3 seed ${seed}: 100 steps in ${(Date.now() - t0) / 1000} s`);
  }
} else if (mode === 'long') {
  // 64^3 Taylor-Green, nu = 1/1600, fixed dt = 0.004 to t = 10
  const s = NS.createSolver({ N: 64, Re: 1600, ic: 'tgv', dt: 0.004 });
  const t0 = Date.now();
  s.run(1e9, 10.0);
  const ser = s.st.series;
  writeJSON('tgv64_dt004_t10.json', { opts: { N: 64, Re: 1600, ic: 'tgv', dt: 0.004 }, st: stScalars(s),
    series: { t: ser.t, E: ser.E, Z: ser.Z, eps: ser.eps, omMax: ser.omMax, uMax: ser.uMax }, secs: (Date.now() - t0) / 1000 });
  console.log('tgv64 dt 0.004 to t=10:', s.st.step, 'steps', (Date.now() - t0) / 1000, 's');
}
