/*
 * Batched ensembles and a continuous-batching request loop over the slot pools of nslab.js.
 *
 * runEnsemble(): many independent solves of uneven N stepped together, one pool per N, every pool's RK4 step a
 *   captured hipGraph, the pools on their own streams so they overlap on the device.
 * serve(): refinement-ladder and Reynolds-sweep requests arrive as a Poisson process; each is admitted (all members
 *   at once) when the pools have free slots or the HBM budget allows a new pool, runs resident on the device beside
 *   the other requests, and completes with its verdict: a ladder the reference's (studySummary, nslab.js:652-666,
 *   over the per-level peaks that nslab-worker.js:15-19 books), a sweep the peaks' scaling with Re (sweepSummary).
 *   Reports p50/p99 completion latency and requests per hour.
 *
 * CLI (one JSON object on stdout):
 *   node serve.js ensemble [--steps 300] [--seq-per-n 8]
 *   node serve.js serve [--requests 60] [--rate 2] [--seed 1] [--chunk 32] [--budget-gb 250] [--sweeps 0] [--cold]
 *                       (--sweeps F: the fraction of requests that are Reynolds sweeps; --cold: no pools created
 *                       before the first arrival)
 */
'use strict';
const path = require('path');
const { performance } = require('perf_hooks');
const NS = require(path.join(__dirname, 'nslab.js'));

// Mulberry32 (the reference's PRNG, nslab.js:161): 32-bit state, one Weyl step and an xor-multiply hash per draw.
function mulberry32(seed) {
  let a = seed | 0;
  return () => {
    a = (a + 0x6D2B79F5) | 0;
    let t = Math.imul(a ^ (a >>> 15), 1 | a);
    t = (t + Math.imul(t ^ (t >>> 7), 61 | t)) ^ t;
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}
const logUniform = (u, lo, hi) => lo * Math.pow(hi / lo, u);
const sleepMs = ms => { if (ms > 0) Atomics.wait(new Int32Array(new SharedArrayBuffer(4)), 0, 0, ms); };
const quantile = (xs, q) => {
  if (!xs.length) return null;
  const a = xs.slice().sort((x, y) => x - y), i = Math.min(a.length - 1, Math.max(0, Math.ceil(q * a.length) - 1));
  return a[i];
};

/** Launch one chunk on every pool with running slots, then wait for all of them (the pools overlap). */
function stepPools(pools, k) {
  const live = pools.filter(p => p.running > 0);
  for (const p of live) p.pool.launch(k);
  for (const p of live) p.running = p.pool.collect();
  return live.length;
}

/** Advance every pool on its own clock: collect a pool whose chunk has finished and relaunch it while it has steps
 *  left, so small grids are not held back by the chunk time of large ones. A pool on hold (a queued request waits to
 *  be admitted into it) is collected but not relaunched. Returns true if a pool was collected. */
function pump(pools, k) {
  let got = false;
  for (const p of pools) {
    if (p.inflight && p.pool.ready()) { p.running = p.pool.collect(); p.inflight = false; got = true; }
    if (!p.inflight && p.running > 0 && !p.hold) { p.pool.launch(k); p.inflight = true; }
  }
  return got;
}

// ------------------------------------------------------------------------------------------------ ensemble
const ENSEMBLE_MIX = [[32, 64], [48, 32], [64, 24], [96, 8]];

/** The batched ensemble: 64x32^3, 32x48^3, 24x64^3, 8x96^3, random IC k0 = 4, E0 = 0.5, seed = solve index + 1,
 *  Re log-uniform in [707, 4000] from Mulberry32(seed), CFL 0.5, `steps` RK4 steps each. */
function ensembleSpec(opts = {}) {
  const mix = opts.mix || ENSEMBLE_MIX, rnd = mulberry32(opts.seed || 20240607), out = [];
  let i = 0;
  for (const [N, count] of mix) for (let c = 0; c < count; c++, i++) {
    out.push({ id: i, N, opts: { N, Re: logUniform(rnd(), 707, 4000), cfl: 0.5, ic: 'random', icParams: { seed: i + 1, k0: 4, E0: 0.5 } } });
  }
  return out;
}

function runEnsemble(opts = {}) {
  const steps = opts.steps || 300, chunk = opts.chunk || 64, spec = ensembleSpec(opts);
  const byN = new Map();
  for (const s of spec) { if (!byN.has(s.N)) byN.set(s.N, []); byN.get(s.N).push(s); }
  const tA = performance.now();
  const pools = [];
  for (const [N, list] of byN) {
    const pool = new NS.Pool(N, list.length), entry = { N, pool, running: 0, members: [] };
    for (const s of list) {
      const solver = pool.admit(s.opts);
      pool.setTarget(solver, steps);
      entry.members.push({ spec: s, solver });
    }
    entry.running = list.length;
    pools.push(entry);
  }
  // opts.group: every pool's step in one captured graph (one launch per ensemble step); else a graph per pool, each
  // on its own stream
  const group = opts.group ? new NS.PoolGroup(pools.map(p => p.pool)) : null;
  const tB = performance.now();
  let chunks = 0;
  if (group) {
    while (pools.some(p => p.running > 0)) {
      group.launch(chunk);
      for (const p of pools) p.running = p.pool.collect();
      chunks++;
    }
  } else {
    while (stepPools(pools, chunk) > 0) chunks++;
  }
  const tC = performance.now();
  if (group) group.destroy();
  const solves = spec.length, stepsDone = pools.reduce((a, p) => a + p.members.reduce((b, m) => b + m.solver.st.step, 0), 0);
  const res = {
    solves, steps, chunks, one_graph: !!group, admit_s: (tB - tA) / 1e3, run_s: (tC - tB) / 1e3,
    solve_steps_per_s: stepsDone / ((tC - tB) / 1e3), ensemble_steps_per_s: steps / ((tC - tB) / 1e3),
    graphed: pools.every(p => p.pool.info().graphed === 1),
    pools: pools.map(p => ({ N: p.N, slots: p.pool.slots, bytes: p.pool.info().bytes })),
    allStepped: pools.every(p => p.members.every(m => m.solver.st.step === steps)),
  };
  // sequential single solves on the same GPU: `seqPerN` solves of each N, run one at a time through createSolver,
  // extrapolated by the mix counts; the first solve of each N is also checked bit for bit against its pool slot
  const seqPerN = opts.seqPerN != null ? opts.seqPerN : 8;
  for (const p of pools) p.pool.destroy();   // the series are on the host; free the device for the sequential leg
  // The steps alone are timed against the pools' run_s; creation and release (initial field, first rhs, device
  // memory) against the pools' admission, in the totals.
  if (seqPerN > 0) {
    let tSeq = 0, tSeqAll = 0, identical = true;
    const perN = [];
    for (const p of pools) {
      const take = p.members.slice(0, Math.min(seqPerN, p.members.length));
      let run = 0, all = 0;
      for (const m of take) {
        const t0 = performance.now();
        const s = NS.createSolver(m.spec.opts);
        const t1 = performance.now();
        s.run(steps);
        const t2 = performance.now();
        if (m === take[0]) {
          const a = s.st.series, b = m.solver.st.series;
          identical = identical && a.E.length === b.E.length && a.E.every((v, i) => v === b.E[i]) &&
            a.omMax.every((v, i) => v === b.omMax[i]) && s.st.t === m.solver.st.t;
        }
        s.destroy();
        run += (t2 - t1) / 1e3;
        all += (performance.now() - t0) / 1e3;
      }
      perN.push({ N: p.N, sampled: take.length, s_per_solve: run / take.length, s_per_solve_with_setup: all / take.length });
      tSeq += run / take.length * p.members.length;
      tSeqAll += all / take.length * p.members.length;
    }
    Object.assign(res, { sequential_s: tSeq, sequential_with_setup_s: tSeqAll, sequential_per_n: perN,
      speedup_vs_sequential: tSeq / res.run_s, speedup_with_setup: tSeqAll / (res.admit_s + res.run_s),
      pool_equals_single: identical });
  }
  return res;
}

// ------------------------------------------------------------------------------------------------ serving
const LADDERS = [[48, 64, 96], [64, 96, 128], [96, 128, 192], [128, 192, 256]];
// Reynolds sweeps: one grid, three consecutive Re of the sqrt(2) series 707 .. 4000 that the reference's tube runs
// use (research/nslab: Re 707, 1000, 1414, 2000)
const SWEEP_GRIDS = [64, 96, 128, 192];
const SWEEP_RE = [707, 1000, 1414, 2000, 2828, 4000];

/** Requests: Poisson arrivals at `rate`/s (Mulberry32(seed)), IC tubes or random (seeded), a step budget of
 *  200..1000 steps, and either a refinement ladder from LADDERS at one Re log-uniform in [707, 4000] (the budget is
 *  the finest rung's), or, with probability `sweepFraction` (default 0), a Reynolds sweep: one grid from SWEEP_GRIDS
 *  at three consecutive Re of SWEEP_RE. Each request lists its members {N, Re}. */
function requestSpec(opts = {}) {
  const n = opts.requests || 60, rate = opts.rate || 2, rnd = mulberry32(opts.seed || 1), out = [];
  const sweepFraction = opts.sweepFraction || 0;
  let t = 0;
  for (let i = 0; i < n; i++) {
    t += -Math.log(1 - rnd()) / rate;
    const sweep = sweepFraction > 0 && rnd() < sweepFraction;
    let q;
    if (sweep) {
      const grids = opts.sweepGrids || SWEEP_GRIDS;
      const N = grids[Math.floor(rnd() * grids.length)], k = Math.floor(rnd() * (SWEEP_RE.length - 2));
      q = { kind: 'sweep', N, Res: SWEEP_RE.slice(k, k + 3) };
      q.members = q.Res.map(Re => ({ N, Re }));
    } else {
      const ladder = (opts.ladders || LADDERS)[Math.floor(rnd() * (opts.ladders || LADDERS).length)];
      q = { kind: 'ladder', ladder };
    }
    const ic = rnd() < 0.5 ? 'tubes' : 'random';
    const Re = logUniform(rnd(), 707, 4000), steps = 200 + Math.floor(rnd() * 801);
    if (!sweep) { q.Re = Re; q.members = q.ladder.map(N => ({ N, Re })); }
    out.push(Object.assign(q, { id: i, arrival: t, ic, steps, icParams: ic === 'random' ? { seed: i + 1, k0: 4, E0: 0.5 } : {} }));
  }
  return out;
}

/** A Reynolds sweep's summary from its levels (one per Re, booked as a ladder's): per Re the peaks and E(t_end), the
 *  least-squares slopes of log peak max|omega| and log peak eps against log Re, and the worst health grade. */
function sweepSummary(levels) {
  const x = levels.map(l => Math.log(l.Re));
  const slope = (y) => {
    const n = x.length, mx = x.reduce((a, b) => a + b, 0) / n, my = y.reduce((a, b) => a + b, 0) / n;
    let sxy = 0, sxx = 0;
    for (let i = 0; i < n; i++) { sxy += (x[i] - mx) * (y[i] - my); sxx += (x[i] - mx) * (x[i] - mx); }
    return sxx > 0 ? sxy / sxx : null;
  };
  const rank = { PASS: 0, WARN: 1, FAIL: 2 };
  const worst = levels.map(l => l.health || 'PASS').reduce((a, b) => (rank[b] > rank[a] ? b : a), 'PASS');
  const sOm = slope(levels.map(l => Math.log(l.peak.omMax))), sEps = slope(levels.map(l => Math.log(l.epsPeak.v)));
  return { Re: levels.map(l => l.Re), omMaxPeak: levels.map(l => l.peak.omMax), epsPeak: levels.map(l => l.epsPeak.v),
    Eend: levels.map(l => l.Eend), slope_omMax: sOm, slope_eps: sEps, verdict: worst,
    text: `${levels[0].label} Re ${levels.map(l => l.Re).join(', ')}: d log ω_max / d log Re ${sOm.toFixed(3)}, ` +
      `d log ε_peak / d log Re ${sEps.toFixed(3)}` };
}

// Slots per pool of grid N: about `gbPerPool` GB of slots, at most 32.
const poolSlots = (N, gbPerPool) => Math.max(1, Math.min(32, Math.floor(gbPerPool * 1e9 / NS.Pool.slotBytes(N))));

/** A request's levels in the form nslab-worker.js:15-19 books them, then the ladder verdict. */
function finishRung(r) {
  const s = r.solver.st.series;
  const peak = (arr) => { let i = 0; for (let k = 0; k < arr.length; k++) if (arr[k] > arr[i]) i = k; return { v: arr[i], t: s.t[i] }; };
  const po = peak(s.omMax);
  return { label: `${r.N}³`, N: r.N, Re: r.Re, peak: { omMax: po.v, t: po.t }, epsPeak: peak(s.eps),
    Eend: s.E[s.E.length - 1], steps: r.solver.st.step, t: r.solver.st.t, health: r.solver.health().worst };
}

function serve(opts = {}) {
  const reqs = requestSpec(opts), chunk = opts.chunk || 32, cfl = 0.5;
  const mem = NS.deviceMemory();
  const budget = Math.min(mem.free * 0.9, (opts.budgetGB || 250) * 1e9);
  const gbPerPool = opts.gbPerPool || 16;
  const pools = [];   // { N, pool, running }
  let reserved = 0, nextArrival = 0, completed = 0;
  const queue = [], results = new Array(reqs.length), live = [];
  // Warm start (opts.warm, default on): one pool per grid the requests can ask for, created before the first arrival,
  // as a server allocates its resident state at start-up; the clock starts after it. More pools are still created
  // on demand while the budget allows.
  const tw0 = performance.now();
  if (opts.warm !== false) {
    const grids = [...new Set(reqs.flatMap(q => q.members.map(m => m.N)))].sort((a, b) => a - b);
    for (const N of grids) {
      const bytes = poolSlots(N, gbPerPool) * NS.Pool.slotBytes(N);
      if (reserved + bytes > budget) break;
      const p = { N, pool: new NS.Pool(N, poolSlots(N, gbPerPool)), running: 0, inflight: false, hold: false };
      reserved += p.pool.info().bytes;
      pools.push(p);
    }
  }
  const warmS = (performance.now() - tw0) / 1e3;
  const t0 = performance.now(), now = () => (performance.now() - t0) / 1e3;

  // free slots of N not yet claimed by the request being admitted; new pools while the budget allows
  function claim(N, claimed) {
    for (const p of pools) if (p.N === N && p.pool.free - (claimed.get(p) || 0) > 0) return p;
    const slots = poolSlots(N, gbPerPool), bytes = slots * NS.Pool.slotBytes(N);
    if (reserved + bytes > budget) return null;
    const p = { N, pool: new NS.Pool(N, slots), running: 0, inflight: false, hold: false };
    reserved += p.pool.info().bytes;
    pools.push(p);
    return p;
  }
  // A member's initial field and first rhs run on its pool's stream, so admitting into a pool with a chunk in flight
  // would wait for that chunk while no other pool is relaunched. A request whose pools are busy puts them on hold
  // instead (they finish their chunk and are not relaunched) and is admitted once they are idle.
  function tryAdmit(q) {
    const claimed = new Map(), at = [];
    for (const m of q.members) {
      const p = claim(m.N, claimed);
      if (!p) return false;
      claimed.set(p, (claimed.get(p) || 0) + 1);
      at.push(p);
    }
    if (at.some(p => p.inflight)) {
      for (const p of at) p.hold = true;
      return false;
    }
    // the last member first (a ladder's finest rung; a sweep's members share grid and IC): its initial max|u| fixes
    // the common end time (the step budget at CFL 0.5 on that grid)
    const rungs = q.members.map((m, i) => ({ N: m.N, Re: m.Re, p: at[i] }));
    const mk = r => r.p.pool.admit({ N: r.N, Re: r.Re, cfl, ic: q.ic, icParams: q.icParams });
    const fine = rungs[rungs.length - 1];
    fine.solver = mk(fine);
    const dt0 = cfl * (2 * Math.PI / fine.N) / Math.max(fine.solver.last.uMax, 1e-9);
    const tEnd = q.steps * dt0;
    for (const r of rungs) {
      if (!r.solver) r.solver = mk(r);
      r.p.pool.setTarget(r.solver, 20 * q.steps, tEnd);
      r.p.running++;
    }
    live.push({ q, rungs, tEnd, admitted: now(), left: rungs.length });
    return true;
  }

  let admitS = 0;   // host time spent admitting (initial fields and first rhs of the new members)
  while (completed < reqs.length) {
    const t = now();
    while (nextArrival < reqs.length && reqs[nextArrival].arrival <= t) queue.push(reqs[nextArrival++]);
    for (const p of pools) p.hold = false;   // set again below by the requests still waiting for busy pools
    for (let i = 0; i < queue.length;) {
      const a0 = now(), ok = tryAdmit(queue[i]);
      admitS += now() - a0;
      if (ok) queue.splice(i, 1); else i++;
    }
    if (!live.length) {
      if (queue.length) throw new Error(`request ${queue[0].id} does not fit the HBM budget`);
      sleepMs(1000 * (reqs[nextArrival].arrival - now()));
      continue;
    }
    if (!pump(pools, chunk)) { sleepMs(0.2); continue; }
    const tDone = now();
    for (let i = 0; i < live.length;) {
      const L = live[i];
      for (const r of L.rungs) {
        if (r.done) continue;
        const st = r.solver.st;
        if (st.t >= L.tEnd - 1e-12 || st.step >= 20 * L.q.steps) {
          r.done = finishRung(r);
          r.p.pool.release(r.solver);
          L.left--;
        }
      }
      if (L.left === 0) {
        const levels = L.rungs.map(r => r.done);
        results[L.q.id] = { id: L.q.id, kind: L.q.kind, ladder: L.q.ladder, N: L.q.N, Res: L.q.Res, ic: L.q.ic,
          Re: L.q.Re, arrival: L.q.arrival, admitted: L.admitted, finished: tDone, latency: tDone - L.q.arrival,
          tEnd: L.tEnd, levels,
          summary: L.q.kind === 'sweep' ? sweepSummary(levels) : NS.studySummary(levels, 'grid') };
        completed++;
        live.splice(i, 1);
      } else i++;
    }
  }
  const wall = now(), lat = results.map(r => r.latency), rungSteps = results.reduce((a, r) => a + r.levels.reduce((b, l) => b + l.steps, 0), 0);
  const solves = results.reduce((a, r) => a + r.levels.length, 0), sweeps = results.filter(r => r.kind === 'sweep').length;
  const poolList = pools.map(p => ({ N: p.N, slots: p.pool.slots }));
  for (const p of pools) p.pool.destroy();
  return { requests: reqs.length, sweeps, rate: opts.rate || 2, wall_s: wall, requests_per_hour: reqs.length / wall * 3600,
    ladders_per_hour: (reqs.length - sweeps) / wall * 3600, sweeps_per_hour: sweeps / wall * 3600,
    rung_solves_per_hour: solves / wall * 3600, p50_latency_s: quantile(lat, 0.5), p99_latency_s: quantile(lat, 0.99),
    rung_steps: rungSteps, pools: poolList, reserved_bytes: reserved, admit_host_s: admitS, warm_start_s: warmS,
    budget_bytes: budget, results };
}

/** Sequential single-solve execution of the same requests (createSolver + run to the request's end time, one rung
 *  at a time), for `sample` requests; returns seconds per request and the rung states' end clocks. */
function serveSequential(reqs, results, sample) {
  const t0 = performance.now(), out = [];
  for (const q of reqs.slice(0, sample)) {
    const R = results[q.id], levels = [];
    for (const m of q.members) {
      const s = NS.createSolver({ N: m.N, Re: m.Re, cfl: 0.5, ic: q.ic, icParams: q.icParams });
      s.run(20 * q.steps, R.tEnd);
      levels.push({ N: m.N, Re: m.Re, steps: s.st.step, t: s.st.t, Eend: s.st.series.E[s.st.series.E.length - 1] });
      s.destroy();
    }
    out.push({ id: q.id, levels });
  }
  return { seconds: (performance.now() - t0) / 1e3, requests: out };
}

/** Single-solve cost on this GPU per grid: seconds to create (IC + first rhs) and seconds per step, from one solve
 *  of each N run alone for `steps` steps after a short warm-up. */
function singleSolveCosts(Ns, steps = 60) {
  const out = {};
  for (const N of Ns) {
    const w = NS.createSolver({ N, Re: 1600, cfl: 0.5, ic: 'random' });
    w.run(5);
    const t0 = performance.now();
    const s = NS.createSolver({ N, Re: 1600, cfl: 0.5, ic: 'random' });
    const t1 = performance.now();
    s.run(steps);
    const t2 = performance.now();
    out[N] = { create_s: (t1 - t0) / 1e3, step_s: (t2 - t1) / 1e3 / steps };
    w.destroy();
    s.destroy();
  }
  return out;
}

/** Sequential single-solve time of served requests from the per-grid costs and each rung's actual step count. */
function sequentialEstimate(results, costs) {
  let t = 0;
  for (const r of results) for (const l of r.levels) t += costs[l.N].create_s + l.steps * costs[l.N].step_s;
  return t;
}

module.exports = { mulberry32, singleSolveCosts, sequentialEstimate, ensembleSpec, runEnsemble, requestSpec, serve,
  serveSequential, sweepSummary, LADDERS, SWEEP_GRIDS, SWEEP_RE, ENSEMBLE_MIX };

if (require.main === module) {
  const argv = process.argv.slice(2), mode = argv[0];
  const arg = (name, d) => { const i = argv.indexOf('--' + name); return i >= 0 ? +argv[i + 1] : d; };
  let out;
  if (mode === 'ensemble') {
    out = runEnsemble({ steps: arg('steps', 300), seqPerN: arg('seq-per-n', 8), chunk: arg('chunk', 64),
      group: argv.includes('--one-graph') });
  } else if (mode === 'serve') {
    const o = { requests: arg('requests', 60), rate: arg('rate', 2), seed: arg('seed', 1), chunk: arg('chunk', 32),
      budgetGB: arg('budget-gb', 250), sweepFraction: arg('sweeps', 0), warm: !argv.includes('--cold') };
    out = serve(o);
    // sequential single-solve execution of the same work: per-grid costs measured here, times each rung's steps;
    // a sample of requests is also run sequentially for real, to check the estimate and bit-identity
    const Ns = [...new Set(LADDERS.flat().concat(o.sweepFraction > 0 ? SWEEP_GRIDS : []))].sort((a, b) => a - b);
    const costs = singleSolveCosts(Ns);
    const seqAll = sequentialEstimate(out.results, costs);
    // the work served: RK4 steps per grid over every rung (the bench prices them with the HBM traffic model)
    const stepsPerGrid = {};
    for (const r of out.results) for (const l of r.levels) stepsPerGrid[l.N] = (stepsPerGrid[l.N] || 0) + l.steps;
    Object.assign(out, { single_solve_costs: costs, sequential_estimate_s: seqAll,
      throughput_vs_sequential: seqAll / out.wall_s, steps_per_grid: stepsPerGrid });
    const seq = arg('seq-sample', 0);
    if (seq > 0) {
      const reqs = requestSpec(o), sq = serveSequential(reqs, out.results, seq);
      // the served end states of the sampled requests must equal the sequential ones bit for bit
      const same = sq.requests.every(r => r.levels.every((l, i) => {
        const m = out.results[r.id].levels[i];
        return l.steps === m.steps && l.t === m.t && l.Eend === m.Eend;
      }));
      const est = sequentialEstimate(out.results.slice(0, seq), costs);
      Object.assign(out, { sequential_sample: seq, sequential_sample_s: sq.seconds, sequential_sample_estimate_s: est,
        served_equals_sequential: same });
    }
    if (!argv.includes('--full')) {
      out.results = out.results.map(r => ({ id: r.id, kind: r.kind, ladder: r.ladder, N: r.N, Res: r.Res,
        latency: r.latency, verdict: r.summary && r.summary.verdict }));
    }
  } else {
    process.stderr.write('usage: node serve.js ensemble|serve [options]\n');
    process.exit(2);
  }
  process.stdout.write(JSON.stringify(out) + '\n');
}
