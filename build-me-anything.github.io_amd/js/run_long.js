#!/usr/bin/env node
/*
 * Long NSLab runs on the GPU with snapshots, checkpoint/resume and continuously written partial results: the batch
 * harness of pocket-wind-tunnel/test/run-ns-long.js and of the GPU runner (gpu/nslab_gpu.py:504-560), with their
 * on-disk formats, so the archive's readers take these runs unchanged.
 *
 *   node run_long.js --N 256 --Re 1600 --tEnd 10 --out <dir> [--ic tgv] [--cfl 0.4] [--dt 0] [--snap 0.5]
 *                    [--ckpt 2] [--chunk 16] [--icParams '{"seed":3}'] [--no-peak-track] [--bench]
 *                    [--compare <archived final.json>] [--fp32]
 *
 * Writes into <dir>: run.log, partial.json (rewritten at every snapshot), slices/t<t>_<kind>_<axis><i>.f32,
 * checkpoint.bin (the six state planes, float64, 6 x N*N*(N/2+1)) + checkpoint.json (atomic rename; a rerun with the
 * same N, Re and IC resumes from them) and final.json. Steps run on the device in chunks of --chunk between host
 * visits; snapshot and checkpoint instants are the first step at or after each multiple of --snap / --ckpt, as in
 * the reference's step-by-step loop. The running grid maximum of |omega| is followed as the GPU runner does
 * (gpu/nslab_gpu.py:551-553): whenever it exceeds the last tracked value by 2 %, the interpolated maximum is
 * evaluated (at the end of that chunk) and appended to peakTrack.
 */
'use strict';
const fs = require('fs');
const path = require('path');
const NS = require(path.join(__dirname, 'nslab.js'));

function parseArgs(argv) {
  const arg = (k, d) => { const i = argv.indexOf('--' + k); return i >= 0 ? argv[i + 1] : d; };
  const o = {
    N: +arg('N', 64), Re: +arg('Re', 1600), tEnd: +arg('tEnd', 10), ic: arg('ic', 'tgv'), cfl: +arg('cfl', 0.4),
    dt: +arg('dt', 0), snap: +arg('snap', 0.5), ckpt: +arg('ckpt', 2), chunk: +arg('chunk', 16),
    icParams: JSON.parse(arg('icParams', '{}')), peakTrack: !argv.includes('--no-peak-track'),
    bench: argv.includes('--bench'), compare: arg('compare', null), fp32: argv.includes('--fp32'),
  };
  o.out = arg('out', path.join(process.cwd(), `${o.ic}-Re${o.Re}-N${o.N}`));
  return o;
}

function runLong(a) {
  fs.mkdirSync(path.join(a.out, 'slices'), { recursive: true });
  const log = s => {
    const line = `[${new Date().toISOString().slice(0, 19).replace('T', ' ')}] ${s}`;
    console.log(line);
    fs.appendFileSync(path.join(a.out, 'run.log'), line + '\n');
  };
  const build = { version: NS.VERSION, library: NS.addonVersion };
  const device = 'MI355X (gfx950), ' + NS.addonVersion;
  // the header ends with the working precision, as the GPU runner's does (nslab_gpu.py:485)
  const precision = a.fp32 ? 'float32' : 'float64';
  log(`NSLab ${NS.VERSION} (${NS.addonVersion}) — ${a.ic} N=${a.N}³ Re=${a.Re} tEnd=${a.tEnd} cfl=${a.cfl} snap=${a.snap} ckpt=${a.ckpt}  node ${process.version}  ${precision}`);
  const t0 = Date.now();
  const s = NS.createSolver({ N: a.N, Re: a.Re, ic: a.ic, cfl: a.cfl, dt: a.dt, icParams: a.icParams, fp32: a.fp32 });
  log(`solver ready in ${((Date.now() - t0) / 1000).toFixed(1)} s, device ${(s.deviceBytes() / 1e9).toFixed(2)} GB, kc ${s.kc}, E0 ${s.st.E0}, Z0 ${s.st.Z0}`);
  // run until the first step with t >= tStop (the reference checks `t >= next - 1e-9` after every step)
  const runTo = (tStop, maxSteps) => {
    let left = maxSteps != null ? maxSteps : Infinity;
    while (s.st.t < tStop - 1e-9 && left > 0) {
      const k = Math.min(a.chunk, left);
      const before = s.st.step;
      s.run(k, tStop - 1e-9 + 1e-12);
      left -= s.st.step - before;
      afterChunk();
      if (s.st.step === before) break;
    }
  };
  let peakTrack = [], peakGrid = 0;
  function afterChunk() {
    if (!a.peakTrack || !(s.last.omMax > 1.02 * peakGrid)) return;
    peakGrid = s.last.omMax;
    const im = s.interpMax(8, 12);
    peakTrack.push({ t: s.st.t, step: s.st.step, omMax: s.last.omMax, omMaxI: im.omMaxI, pos: im.omMaxIpos });
  }

  if (a.bench) {
    s.run(1);
    const tb = Date.now(); s.run(5); const ms = (Date.now() - tb) / 5;
    const td = Date.now(); s.diagnose();
    log(`bench: ${ms.toFixed(1)} ms/step, diagnose ${Date.now() - td} ms, dt ${s.st.dt.toExponential(3)} → ~${Math.round(a.tEnd / s.st.dt)} steps ≈ ${(a.tEnd / s.st.dt * ms / 3600e3).toFixed(3)} h, device ${(s.deviceBytes() / 1e9).toFixed(2)} GB`);
    return { bench: ms };
  }
  if (a.compare) {
    const ref = JSON.parse(fs.readFileSync(a.compare, 'utf8')), rs = ref.series;
    runTo(ref.t, rs.t.length - 1);
    const ss = s.st.series, n = Math.min(rs.t.length, ss.t.length), worst = {};
    for (const k of ['t', 'E', 'Z', 'eps', 'omMax', 'Pspec']) {
      let w = 0; const scale = Math.max(...rs[k].slice(0, n).map(Math.abs)) || 1;
      for (let i = 0; i < n; i++) w = Math.max(w, Math.abs(ss[k][i] - rs[k][i]) / scale);
      worst[k] = w;
    }
    const d = s.diagnose(), h = s.health();
    log(`compare vs ${path.basename(a.compare)} over ${n} steps: max rel diff ` + Object.entries(worst).map(([k, v]) => `${k} ${v.toExponential(2)}`).join(' ') +
      `; final E ${s.last.E.toPrecision(12)} (ref ${rs.E[n - 1].toPrecision(12)}); Pspec ${d.Pspec.toExponential(6)} Pphys ${d.Pphys.toExponential(6)}; divMax ${d.divMax.toExponential(1)}; ebal ${s.st.maxEbal.toExponential(1)}; health ${h.worst}`);
    const ex = s.exactError();
    if (ex) log(`exact-solution L∞ error ${ex.linf.toExponential(2)}`);
    return { compare: worst, steps: n };
  }

  // ---- resume ----
  const ckBin = path.join(a.out, 'checkpoint.bin'), ckJson = path.join(a.out, 'checkpoint.json');
  let outs = [], nextSnap = a.snap, nextCkpt = a.ckpt, elapsedBefore = 0;
  if (fs.existsSync(ckBin) && fs.existsSync(ckJson)) {
    const meta = JSON.parse(fs.readFileSync(ckJson, 'utf8'));
    if (meta.N === a.N && meta.Re === a.Re && meta.ic === a.ic) {
      const buf = fs.readFileSync(ckBin), S = s._S, n = S[0].length;
      for (let c = 0; c < 6; c++) S[c].set(new Float64Array(buf.buffer, buf.byteOffset + c * n * 8, n));
      Object.assign(s.st, { t: meta.t, step: meta.step, series: meta.series, maxEbal: meta.maxEbal, maxZbal: meta.maxZbal,
        maxTnl: meta.maxTnl, divMax: meta.divMax, E0: meta.E0, Z0: meta.Z0 });
      s.st.outputs.length = 0;
      for (const o of meta.outs) s.st.outputs.push(o);
      outs = meta.outs.slice(); nextSnap = meta.nextSnap; nextCkpt = meta.nextCkpt; elapsedBefore = meta.elapsed || 0;
      peakTrack = meta.peakTrack || []; peakGrid = meta.peakGrid || 0;
      log(`resumed from checkpoint at t=${meta.t.toFixed(4)} step ${meta.step} (${outs.length} snapshots)`);
    } else log('checkpoint present but for a different case — ignored');
  }
  const health = () => {
    const h = s.health();
    const o = { worst: h.worst, worstEnd: h.worstEnd, rows: h.rows, note: h.note };
    if (h.precision) o.precision = h.precision;
    return o;
  };
  function writePartial(final) {
    const j = { instrument: 'Pocket Wind Tunnel NSLab ' + NS.VERSION + ' (MI355X)', build, device, precision,
      case: { ic: a.ic, N: a.N, Re: a.Re, nu: 1 / a.Re, cfl: a.cfl, tEnd: a.tEnd, snapEvery: a.snap, icParams: a.icParams },
      t: s.st.t, steps: s.st.step, elapsed_s: elapsedBefore + (Date.now() - t0) / 1000, series: s.st.series,
      snapshots: outs, peakTrack, health: health(), final: !!final,
      disclaimer: 'Numerical evidence only; nothing here proves regularity or blow-up of the Navier–Stokes equations.' };
    const f = path.join(a.out, final ? 'final.json' : 'partial.json');
    fs.writeFileSync(f + '.tmp', JSON.stringify(j));
    fs.renameSync(f + '.tmp', f);
  }
  function snapshot() {
    const d = s.diagnose(); outs.push(d);
    const tag = d.t.toFixed(2).replace('.', 'p');
    for (const kind of ['vort', 'q', 'stretch'])
      for (const [axis, idx] of [['z', 0], ['z', a.N / 4], ['x', a.N / 4], ['x', a.N / 2]])
        fs.writeFileSync(path.join(a.out, 'slices', `t${tag}_${kind}_${axis}${idx}.f32`), Buffer.from(s.slice(kind, axis, idx).buffer));
    writePartial(false);
    const h = health(), done = s.st.t / a.tEnd, el = elapsedBefore + (Date.now() - t0) / 1000;
    const interp = d.omMaxI != null ? ` (interp ${d.omMaxI.toFixed(3)})` : '';
    log(`t ${d.t.toFixed(3)} step ${s.st.step} dt ${s.st.dt.toExponential(2)}  E ${d.E.toFixed(6)} Z ${d.Z.toFixed(4)} ε ${d.eps.toFixed(6)} max|ω| ${d.omMax.toFixed(3)}${interp} ⟨ωSω⟩ ${d.Pspec.toFixed(4)} (phys ${d.Pphys.toFixed(4)}) kmaxη ${d.kmaxEta.toFixed(2)} S ${d.skewIso.toFixed(3)} align ${d.align.map(x => x.toFixed(2)).join('/')} zgap ${d.imageGap.toFixed(2)}  health ${h.worst} (end ${h.worstEnd})  ebal ${s.st.maxEbal.toExponential(1)} zbal ${s.st.maxZbal.toExponential(1)}  ${(el / 3600).toFixed(2)} h, ETA ${((el / Math.max(done, 1e-6)) * (1 - done) / 3600).toFixed(2)} h`);
  }
  function checkpoint() {
    const S = s._S, n = S[0].length, buf = Buffer.alloc(6 * n * 8);
    for (let c = 0; c < 6; c++) Buffer.from(S[c].buffer, S[c].byteOffset, n * 8).copy(buf, c * n * 8);
    fs.writeFileSync(ckBin + '.tmp', buf); fs.renameSync(ckBin + '.tmp', ckBin);
    const meta = { N: a.N, Re: a.Re, ic: a.ic, t: s.st.t, step: s.st.step, series: s.st.series, outs, peakTrack, peakGrid,
      nextSnap, nextCkpt, maxEbal: s.st.maxEbal, maxZbal: s.st.maxZbal, maxTnl: s.st.maxTnl, divMax: s.st.divMax,
      E0: s.st.E0, Z0: s.st.Z0, elapsed: elapsedBefore + (Date.now() - t0) / 1000 };
    fs.writeFileSync(ckJson + '.tmp', JSON.stringify(meta)); fs.renameSync(ckJson + '.tmp', ckJson);
    log(`checkpoint written at t=${s.st.t.toFixed(4)}`);
  }
  if (s.st.step === 0) snapshot();
  while (s.st.t < a.tEnd - 1e-12) {
    runTo(Math.min(nextSnap, nextCkpt, a.tEnd - 1e-12 + 1e-9));   // last stop: the step that reaches tEnd - 1e-12
    if (s.st.t >= nextSnap - 1e-9) { nextSnap += a.snap; snapshot(); }
    if (s.st.t >= nextCkpt - 1e-9) { nextCkpt += a.ckpt; checkpoint(); }
  }
  if (outs.length === 0 || outs[outs.length - 1].t < s.st.t - 1e-9) snapshot();
  writePartial(true);
  const ser = s.st.series;
  let ip = 0, io = 0;
  for (let i = 0; i < ser.t.length; i++) { if (ser.eps[i] > ser.eps[ip]) ip = i; if (ser.omMax[i] > ser.omMax[io]) io = i; }
  const pts = peakTrack.concat(outs.filter(o => o.omMaxI != null).map(o => ({ t: o.t, omMaxI: o.omMaxI })));
  const pI = pts.length ? pts.reduce((x, y) => (y.omMaxI > x.omMaxI ? y : x)) : null;
  log(`DONE: ε_max ${ser.eps[ip].toFixed(6)} at t ${ser.t[ip].toFixed(3)}; max|ω| peak ${ser.omMax[io].toFixed(3)} at t ${ser.t[io].toFixed(3)}; interpolated peak ${pI ? pI.omMaxI.toFixed(3) : 'NaN'} at t ${pI ? pI.t.toFixed(3) : 'NaN'}; E(${a.tEnd}) ${ser.E[ser.E.length - 1].toFixed(6)}; health ${health().worst}; ${((elapsedBefore + (Date.now() - t0) / 1000) / 3600).toFixed(2)} h`);
  return { t: s.st.t, steps: s.st.step };
}

module.exports = { runLong, parseArgs };

if (require.main === module) runLong(parseArgs(process.argv.slice(2)));
