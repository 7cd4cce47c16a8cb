#!/usr/bin/env node
/*
 * Grade a fresh run against an archived one before trusting a new machine with a long run: the gate of
 * cloud/parity-check.py:22-59 over the same final.json format. Both runs must be the same case (N, Re, IC, CFL,
 * precision); the series are compared step for step over their overlap, relative to the archived series' largest
 * magnitude. Per-quantity tolerances (deterministic window, before the t ~ 8.5 burst of the tube/TGV flows):
 * t 1e-12, E Z eps 1e-9, max|omega| Pspec 1e-8 (above: WARN); above 1e-6 the gate FAILs.
 *
 *   node parity_check.js <fresh final.json> <archived final.json>      exit 0 PASS/WARN, 1 FAIL, 2 usage
 */
'use strict';
const fs = require('fs');

const KEYS = [['t', 1e-12], ['E', 1e-9], ['Z', 1e-9], ['eps', 1e-9], ['omMax', 1e-8], ['Pspec', 1e-8]];
const FAIL = 1e-6;

function parityCheck(a, b, print = console.log) {
  for (const [f, what] of [['N', 'grid'], ['Re', 'Reynolds number'], ['ic', 'initial condition'], ['cfl', 'CFL']]) {
    if (a.case[f] !== b.case[f]) { print(`PARITY FAIL: ${what} differs - fresh ${a.case[f]} vs archived ${b.case[f]}`); return { verdict: 'FAIL' }; }
  }
  if (a.precision !== b.precision) { print(`PARITY FAIL: precision differs - fresh ${a.precision} vs archived ${b.precision}`); return { verdict: 'FAIL' }; }
  const sa = a.series, sb = b.series, n = Math.min(sa.t.length, sb.t.length);
  if (n < 10) { print(`PARITY FAIL: only ${n} common steps`); return { verdict: 'FAIL' }; }
  if (sa.t[n - 1] > 8.0) print(`PARITY WARN: overlap reaches t = ${sa.t[n - 1].toFixed(2)}; late-time disagreement past the t ~ 8.5 burst is physics`);
  const worst = {}, bad = [], warn = [];
  for (const [k, tol] of KEYS) {
    let scale = 0;
    for (let i = 0; i < n; i++) scale = Math.max(scale, Math.abs(sb[k][i]));
    scale = scale || 1;
    let w = 0;
    for (let i = 0; i < n; i++) w = Math.max(w, Math.abs(sa[k][i] - sb[k][i]) / scale);
    worst[k] = w;
    if (w > FAIL) bad.push(k); else if (w > tol) warn.push(k);
  }
  print(`parity over ${n} steps to t = ${sa.t[n - 1].toFixed(4)}  (${a.device} vs archived ${b.device})`);
  for (const [k, tol] of KEYS) print(`  ${k.padEnd(7)} max rel diff ${worst[k].toExponential(3)}   tol ${tol.toExponential(0)}   ${worst[k] > FAIL ? 'FAIL' : worst[k] > tol ? 'warn' : 'ok'}`);
  const verdict = bad.length ? 'FAIL' : warn.length ? 'WARN' : 'PASS';
  print(verdict === 'FAIL' ? `PARITY FAIL: ${bad.join(', ')}` : verdict === 'WARN' ? `PARITY WARN: ${warn.join(', ')}` : 'PARITY PASS - this machine reproduces the archive.');
  return { verdict, worst, steps: n };
}

module.exports = { parityCheck };

if (require.main === module) {
  const [f, g] = process.argv.slice(2);
  if (!f || !g) { process.stderr.write('usage: node parity_check.js <fresh final.json> <archived final.json>\n'); process.exit(2); }
  const r = parityCheck(JSON.parse(fs.readFileSync(f, 'utf8')), JSON.parse(fs.readFileSync(g, 'utf8')));
  process.exit(r.verdict === 'FAIL' ? 1 : 0);
}
