// Pools and the request loop of the Node host (build-me-anything.github.io_amd/js/serve.js), run under Node on a GPU
// box. Every check is bit-identity: a solve stepped in a pool, or a ladder served beside other requests, equals the
// same solve run alone through createSolver()/run(). Exit status 0 and "ALL PASS" on success.
'use strict';
const path = require('path');
const pkg = path.join(__dirname, '..', '..', 'build-me-anything.github.io_amd', 'js');
const NS = require(path.join(pkg, 'nslab.js'));
const SV = require(path.join(pkg, 'serve.js'));

let fails = 0;
function check(name, ok, detail) {
  if (!ok) fails++;
  console.log(`${ok ? 'PASS' : 'FAIL'}  ${name}${detail != null ? '   ' + detail : ''}`);
}
const sameSeries = (a, b) => ['t', 'E', 'Z', 'eps', 'omMax', 'uMax', 'dt', 'ebal', 'zbal', 'Pspec', 'Pal']
  .every(k => a.st.series[k].length === b.st.series[k].length && a.st.series[k].every((v, i) => v === b.st.series[k][i]));

// 1. a pool stepped asynchronously with members joining and leaving
{
  const pool = new NS.Pool(48, 3);
  check('pool step is a captured graph', pool.info().graphed === 1);
  const specs = [{ Re: 900, cfl: 0.5, ic: 'random', icParams: { seed: 4 } }, { Re: 2000, cfl: 0.5, ic: 'tubes' },
    { Re: 1600, cfl: 0.4, ic: 'tgv' }];
  const steps = [37, 20, 45];
  const m = specs.map((o, i) => { const s = pool.admit(o); pool.setTarget(s, steps[i]); return s; });
  let running = 3, launches = 0;
  while (running > 0) { pool.launch(8); while (!pool.ready()) { /* spin */ } running = pool.collect(); launches++; }
  check('pool drains in ceil(45/8) launches', launches === 6, launches);
  pool.release(m[1]);
  const late = pool.admit({ Re: 700, cfl: 0.5, ic: 'abc' });
  pool.setTarget(late, 11);
  while (pool.launch(4), pool.collect() > 0) { /* until drained */ }
  specs.forEach((o, i) => {
    const s = NS.createSolver(Object.assign({ N: 48 }, o));
    s.run(steps[i]);
    check(`pool member ${i} equals its single solve`, sameSeries(s, m[i]) && s.st.t === m[i].st.t);
  });
  const s = NS.createSolver({ N: 48, Re: 700, cfl: 0.5, ic: 'abc' });
  s.run(11);
  check('member admitted into a released slot equals its single solve', sameSeries(s, late));
  check('pool slot health report runs', ['PASS', 'WARN', 'FAIL'].includes(late.health().worst));
}

// 2. the ensemble runner on a small mix
{
  const r = SV.runEnsemble({ mix: [[32, 6], [48, 3]], steps: 40, seqPerN: 2, chunk: 16 });
  check('ensemble: every solve stepped', r.allStepped && r.solves === 9);
  check('ensemble: pool slot equals the single solve', r.pool_equals_single);
  // every pool's step in one captured graph: the same bits as a graph per pool
  const g = SV.runEnsemble({ mix: [[32, 6], [48, 3], [64, 2]], steps: 40, seqPerN: 2, chunk: 16, group: true });
  check('ensemble in one graph: every solve stepped', g.allStepped && g.solves === 11 && g.one_graph);
  check('ensemble in one graph: pool slot equals the single solve', g.pool_equals_single);
}

// 2b. _S of a pool member follows the device: read, step through the pool, read again
{
  const pool = new NS.Pool(32, 2);
  const a = pool.admit({ Re: 800, ic: 'random', cfl: 0.5, icParams: { seed: 9 } });
  const before = Float64Array.from(a._S[0]);
  pool.setTarget(a, 7);
  while (true) { pool.launch(4); if (pool.collect() === 0) break; }
  const after = a._S[0];
  const ref = NS.createSolver({ N: 32, Re: 800, ic: 'random', cfl: 0.5, icParams: { seed: 9 } });
  ref.run(7);
  check('pool member _S after stepping is the stepped state', after.some((v, i) => v !== before[i]) &&
    after.every((v, i) => v === ref._S[0][i]) && a.st.step === 7);
  ref.destroy();
  pool.destroy();
}

// 3. the request loop: a request served alone and in a mix gives the same levels, equal to sequential solves
{
  const base = { requests: 5, rate: 40, seed: 7, chunk: 16, ladders: [[32, 48, 64], [48, 64, 96]] };
  const mix = SV.serve(base);
  check('serve: all requests completed', mix.results.length === 5 && mix.results.every(r => r && r.levels.length === 3));
  check('serve: latencies reported', mix.p50_latency_s > 0 && mix.p99_latency_s >= mix.p50_latency_s);
  const reqs = SV.requestSpec(base);
  const seq = SV.serveSequential(reqs, mix.results, 5);
  const same = seq.requests.every(q => q.levels.every((l, i) => {
    const m = mix.results[q.id].levels[i];
    return l.steps === m.steps && l.t === m.t && l.Eend === m.Eend;
  }));
  check('serve: every level equals the sequential single solve', same);
  // request 2 alone: requestSpec with one request reproduces only request 0, so serve the same work directly
  const one = SV.serve(Object.assign({}, base, { requests: 1 }));
  const a = one.results[0].levels, b = mix.results[0].levels;
  check('serve: request 0 alone equals request 0 in the mix',
    a.every((l, i) => l.steps === b[i].steps && l.Eend === b[i].Eend && l.peak.omMax === b[i].peak.omMax && l.epsPeak.v === b[i].epsPeak.v));
  check('serve: ladder verdicts present', mix.results.every(r => r.summary && ['PASS', 'WARN', 'FAIL'].includes(r.summary.verdict)));
}

// 4. Reynolds sweeps beside ladders: every member equals its sequential single solve, and a sweep request served
//    alone gives the same levels as in the mix
{
  const base = { requests: 6, rate: 40, seed: 11, chunk: 16, ladders: [[32, 48, 64]], sweepFraction: 0.5,
    sweepGrids: [32, 48] };
  const reqs = SV.requestSpec(base), mix = SV.serve(base);
  const sweeps = mix.results.filter(r => r.kind === 'sweep');
  check('sweeps: the mix holds sweeps and ladders', sweeps.length > 0 && sweeps.length < 6, sweeps.length);
  check('sweeps: three Re per sweep, one grid', sweeps.every(r => r.levels.length === 3 &&
    r.levels.every(l => l.N === r.N) && r.levels.map(l => l.Re).join() === r.Res.join()));
  const seq = SV.serveSequential(reqs, mix.results, 6);
  check('sweeps: every member equals the sequential single solve', seq.requests.every(q => q.levels.every((l, i) => {
    const m = mix.results[q.id].levels[i];
    return l.steps === m.steps && l.t === m.t && l.Eend === m.Eend;
  })));
  check('sweeps: summary has the Re scaling', sweeps.every(r => r.summary && Number.isFinite(r.summary.slope_omMax) &&
    ['PASS', 'WARN', 'FAIL'].includes(r.summary.verdict)));
  const first = reqs.findIndex(q => q.kind === 'sweep');
  const alone = SV.serve(Object.assign({}, base, { requests: first + 1 }));
  const a = alone.results[first].levels, b = mix.results[first].levels;
  check('sweeps: a sweep alone equals the same sweep in the mix',
    a.every((l, i) => l.steps === b[i].steps && l.Eend === b[i].Eend && l.peak.omMax === b[i].peak.omMax));
}

// 5. poolDestroy frees a pool's device memory at once; the handle is then refused
{
  const addon = require(path.join(pkg, '..', 'lib', 'nslab_napi.node'));
  const slotBytes = addon.poolSlotBytes(128);
  const before = NS.deviceMemory().free;
  const p = addon.poolCreate(128, 8);
  const held = before - NS.deviceMemory().free;
  addon.poolAdmit(p, 128, 0, 1 / 1600, 0.4, 1e-3, 0, 0, 0, 0);
  addon.poolDestroy(p);
  const back = NS.deviceMemory().free;
  check('poolDestroy: the pool held its slots', held >= 8 * slotBytes * 0.99, `${held} B for 8 x ${slotBytes} B`);
  check('poolDestroy: memory returned at once', before - back < slotBytes, `${before - back} B still held`);
  let threw = false;
  try { addon.poolInfo(p); } catch (e) { threw = /not a pool handle/.test(e.message); }
  check('poolDestroy: a destroyed pool handle is refused', threw);
}

console.log(fails ? `${fails} FAILED` : 'ALL PASS');
process.exit(fails ? 1 : 0);
