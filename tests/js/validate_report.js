// The host report module (js/report.js) against health reports and a ladder verdict written by the reference itself
// (tests/golden/, made by make_golden.js from nslab.js): every row's label, text and grade, the overall verdicts and
// the summary text must be the reference's, character for character. No GPU: the inputs are the numbers the
// reference's own run produced (state scalars, last rhs, diagnose snapshot); the pile-up rule is the library's.
'use strict';
const path = require('path');
const fs = require('fs');
const report = require(path.join(__dirname, '..', '..', 'build-me-anything.github.io_amd', 'js', 'report.js'));
const addon = require(path.join(__dirname, '..', '..', 'build-me-anything.github.io_amd', 'lib', 'nslab_napi.node'));

const G = name => JSON.parse(fs.readFileSync(path.join(__dirname, '..', 'golden', name), 'utf8'));
let fails = 0;
function same(what, got, want) {
  const a = JSON.stringify(got), b = JSON.stringify(want);
  if (a !== b) { fails++; console.log('FAIL', what, '\n  got ', a, '\n  want', b); } else console.log('ok  ', what);
}

// A finished run as report.js reads it: state scalars after the run, the snapshot diagnose() archived (if any).
function runOf(N, Re, dtFixed, st, snap, withSnapshot) {
  const kc = Math.floor(N / 3), nu = 1 / Re;
  const stx = Object.assign({}, st, { outputs: withSnapshot ? [snap] : [] });
  if (withSnapshot) stx.divMax = Math.max(st.divMax, snap.divMax);
  return { N, kc, nu, Re, cfl: 0.4, dtFixed, st: stx, last: snap, spectrum: () => snap.spectrum,
    pileUp: spec => addon.cutoffPileUp(kc, Float64Array.from(spec)) };
}
function checkHealth(what, run, want) {
  const h = report.healthReport(run, undefined);
  same(what + ': rows', h.rows, want.rows);
  same(what + ': verdicts', [h.worst, h.worstEnd], [want.worst, want.worstEnd]);
  same(what + ': note', h.note, want.note);
}

const tgv = G('tgv32_dt01.json');
checkHealth('TGV 32^3 dt 0.01, 200 steps, no snapshot', runOf(32, 1600, 0.01, tgv.st200, tgv.diag, false), tgv.healthNoDiag);
checkHealth('TGV 32^3 dt 0.01, 200 steps, after diagnose', runOf(32, 1600, 0.01, tgv.st200, tgv.diag, true), tgv.health);
const cfl = G('tgv32_cfl04_t2.json');
checkHealth('TGV 32^3 CFL 0.4 to t 2', runOf(32, 1600, 0, cfl.st, cfl.diag, true), cfl.health);
const tb = G('tubes.json').re500;
checkHealth('tubes 24^3 Re 500, 20 steps', runOf(24, 500, 0, tb.st, tb.diag, true), tb.health);
const ab = G('exact.json').abc;
checkHealth('ABC 16^3 Re 10 to t 1', runOf(16, 10, 0, ab.st, ab.diag, true), ab.health);
const t48 = G('diag_misc.json').tgv48;
checkHealth('TGV 48^3, 20 steps', runOf(48, 1600, 0, t48.st, t48.diag, true), t48.health);

// a study row of each kind and a worst-instant pile-up row (archived snapshots with pile-up > 1)
{
  const run = runOf(32, 1600, 0.01, tgv.st200, tgv.diag, true);
  const s1 = Object.assign({}, tgv.diag, { t: 1.5, pileUp: 1.7 }), s2 = Object.assign({}, tgv.diag, { t: 2.5, pileUp: 1.31 });
  run.st.outputs = [s1, s2, tgv.diag];
  const h = report.healthReport(run, { grid: { text: 'g', verdict: 'WARN' }, time: { text: 't', verdict: 'PASS' } });
  same('study rows', h.rows.slice(17, 19), [['grid convergence', 'g', 'WARN'], ['time-step convergence', 't', 'PASS']]);
  same('worst-instant pile-up row', h.rows[h.rows.length - 1], ['worst instant: cutoff pile-up', '1.70 at t 1.50', 'WARN']);
  const atLeastWarn = v => (v === 'PASS' ? 'WARN' : v);   // a WARN row can only make the verdict worse
  same('verdicts with a WARN study', [h.worst, h.worstEnd], [atLeastWarn(tgv.health.worst), atLeastWarn(tgv.health.worstEnd)]);
}

// the refinement verdict of a 16/24/32 ladder, from the reference's own per-level peaks
const lad = G('ladder_re100.json');
same('studySummary 16/24/32', report.studySummary(lad.levels, 'grid'), lad.summary);
same('studySummary of one level', report.studySummary(lad.levels.slice(0, 1), 'grid'), null);

// the Node host's pile-up rule (the library's, through the addon) on the reference's archived FP64 spectra; the one
// snapshot only the old runner's fixed 1e-20 guard made null is defined under the current rule
for (const c of G('pileup_fp64.json').cases) {
  const k8 = Math.round(0.8 * c.kc), sp = c.spectrum, peak = Math.max(...sp.slice(1));
  const oldGuardOnly = c.pileUp === null && sp[k8] <= 1e-20 && sp[k8] > 100 * Number.EPSILON ** 2 * peak;
  const got = addon.cutoffPileUp(c.kc, Float64Array.from(sp));
  if (oldGuardOnly) same(`pile-up ${c.N}^3 t ${c.t.toFixed(2)} defined`, got !== null, true);
  else same(`pile-up ${c.N}^3 t ${c.t.toFixed(2)}`, got, c.pileUp);
}

// FP32 solves are graded as the GPU runner grades its FP32 mode (nslab_gpu.py:410-459): the strings come from the
// runner's source (tests/golden/formats.json, fp32_grading, extracted by make_format_fixture.py)
const F = G('formats.json').fp32_grading;
same('FP32 row-label suffix is the runner\'s', report.ROUNDOFF, F.label_suffix);
same('FP32 note text is the runner\'s', report.NOTE_FP32, F.note_fp32);
{
  const r64 = runOf(32, 1600, 0.01, tgv.st200, tgv.diag, true), h64 = report.healthReport(r64, undefined);
  const r32 = Object.assign(runOf(32, 1600, 0.01, tgv.st200, tgv.diag, true), { fp32: true });
  // roundoff-level values an FP32 run produces: graded FAIL in FP64, ungraded in FP32
  r32.st.divMax = 3.8e-6; r32.st.maxTnl = 5.7e-6;
  const h32 = report.healthReport(r32, undefined);
  const want = h64.rows.map(r => r.slice());
  want[3] = ['divergence L∞' + F.label_suffix, '3.80e-6', 'N/A'];
  want[4] = ['nonlinear energy transfer |T|/ε' + F.label_suffix, '5.70e-6', 'N/A'];
  same('FP32 health rows: the two roundoff rows ungraded, every other row as in FP64', h32.rows, want);
  same('FP32 verdicts ignore the ungraded rows', [h32.worst, h32.worstEnd], [h64.worst, h64.worstEnd]);
  same('FP32 report: precision and note', [h32.precision, h32.note], [F.precision.fp32, report.NOTE + F.note_fp32]);
  same('FP64 report carries neither', [h64.precision, h64.note], [undefined, report.NOTE]);
  // the same roundoff values in an FP64 run fail
  const r64b = runOf(32, 1600, 0.01, tgv.st200, tgv.diag, true);
  r64b.st.divMax = 3.8e-6; r64b.st.maxTnl = 5.7e-6;
  const hb = report.healthReport(r64b, undefined);
  same('FP64 grades the same values FAIL', [hb.rows[3][2], hb.rows[4][2], hb.worstEnd], ['FAIL', 'FAIL', 'FAIL']);
}
// the host's FP32 pile-up guard (the library's, with the runner's FP32 epsilon) on the archived 640^3 FP32 run:
// ungraded through t = 5.0, graded from t = 5.5 on (gpu/README.md:44-58)
{
  const d = G('pileup_fp32.json');
  const graded = [];
  for (const q of d.snapshots) {
    const sp = new Float64Array(q.nb);
    sp[1] = q.peak;
    sp.set(q.band, d.k8);
    const v = addon.cutoffPileUp(d.kc, sp, F.eps32);
    if (v !== null) { graded.push(+q.t.toFixed(1)); if (v !== q.pileUp) same(`FP32 pile-up t ${q.t}`, v, q.pileUp); }
    if (addon.cutoffPileUp(d.kc, sp) !== q.pileUp) same(`FP64 pile-up t ${q.t}`, addon.cutoffPileUp(d.kc, sp), q.pileUp);
  }
  same('FP32 guard on the archived 640^3 run: first graded snapshot, count', [graded[0], graded.length], [5.5, d.snapshots.length - 11]);
}

console.log(fails ? `${fails} FAILED` : 'ALL PASS');
process.exit(fails ? 1 : 0);
