/*
 * Health report, evidence dossier and refinement verdict of an NSLab solve, as text the reference's callers read
 * (pocket-wind-tunnel/src/nslab.js:595-666: the row labels, thresholds, number formats and verdict strings are the
 * reference's; the code that produces them is this module's own). No GPU and no addon: everything here works on
 * numbers the solver has already produced.
 *
 * A health check is a row {label, value, format, limits, higher}: `value` reads a number (or null: shown as '—',
 * ungraded) from the run context, `limits` = [good, warn] grade it PASS/WARN/FAIL (lower is better unless `higher`).
 * Rows without limits are informational. The worst-instant checks scan the archived snapshots for the least healthy
 * one of each kind.
 */
'use strict';

const TWO_PI = 2 * Math.PI;
const DASH = '—';

const NOTE = 'A run that does not PASS resolution, budgets and refinement cannot support any claim about the ' +
  'equations. Numerical growth ≠ mathematical singularity. The verdict includes the worst archived snapshot; ' +
  'worstEnd is the end-of-run verdict alone.';
// FP32 solves (the GPU runner's grading, nslab_gpu.py:410-459): the rows that measure quantities exactly zero in
// exact arithmetic (divergence after projection, net nonlinear transfer) are roundoff-limited in float32 and
// reported ungraded; the note says what a float32 PASS does not cover.
const ROUNDOFF = ' — float32: roundoff-limited, ungraded';
const NOTE_FP32 = ' float32 run: the divergence and nonlinear-transfer rows are roundoff-limited and reported N/A ' +
  'rather than graded, and the cutoff pile-up is ungraded whenever E(0.8·kc) sits on the fp32 noise floor. A PASS ' +
  'here is therefore weaker than a float64 PASS — it is silent about those quantities, not clear on them. float32 ' +
  'remains exploration grade (the expl- prefix); archived evidence is float64.';
const EQUATIONS = '3D incompressible Navier–Stokes, periodic [0,2π]³, Re = 1/ν, Fourier pseudo-spectral ' +
  '(2/3 dealiased, rotational form), RK4';
const DISCLAIMER = 'Numerical evidence only. Nothing here is a proof of global regularity or of finite-time blow-up ' +
  'of the Navier–Stokes equations; every feature must survive spatial and temporal refinement before it is even a ' +
  'conjecture.';

const fixed = d => v => v.toFixed(d);
const expo = d => v => v.toExponential(d);

/** PASS / WARN / FAIL of v against [good, warn]; `higher`: larger values are healthier. */
function grade(v, limits, higher) {
  const [good, warn] = limits;
  const ok = higher ? x => v >= x : x => v <= x;
  return ok(good) ? 'PASS' : ok(warn) ? 'WARN' : 'FAIL';
}

/** The overall verdict of graded rows: the worst grade present, PASS when none is graded. */
function verdictOf(rows) {
  const rank = { PASS: 0, WARN: 1, FAIL: 2 };
  const worst = rows.reduce((w, r) => Math.max(w, r[2] in rank ? rank[r[2]] : 0), 0);
  return ['PASS', 'WARN', 'FAIL'][worst];
}

/** Index of the largest bin (the first one on ties), counting bin 0. */
function peakBin(spec) {
  let best = 0;
  spec.forEach((v, k) => { if (v > spec[best]) best = k; });
  return best;
}
/** E(kc)/E(peak): how much energy is left at the cutoff. */
const tailRatio = (spec, kc) => spec[kc] / Math.max(spec[peakBin(spec)], 1e-300);
/** Relative gap between the spectral and the physical-space stretching of one snapshot. */
const stretchGap = (q, nu) => Math.abs(q.Pspec - q.Pphys) / Math.max(Math.abs(q.Pspec), 2 * nu * q.P, 1e-300);

// The end-of-run checks, in the reference's row order. ctx: {N, kc, nu, dtFixed, st, d (last), o (last snapshot or
// null), spec, pile, study, fp32}. `roundoff`: ungraded ('N/A', label suffixed) in an FP32 solve.
const END_CHECKS = [
  { label: 'grid', text: c => `${c.N}³ (kmax ${c.kc}, dealiased 2/3)` },
  { label: 'Δt', text: c => c.st.dt.toExponential(3) + (c.dtFixed ? ' fixed' : ' adaptive') },
  { label: 'CFL (u_max Δt/Δx)', value: c => c.d.uMax * c.st.dt / (TWO_PI / c.N), format: fixed(3), limits: [0.8, 1.2] },
  { label: 'divergence L∞', value: c => c.st.divMax, format: expo(2), limits: [1e-10, 1e-6], roundoff: true },
  { label: 'nonlinear energy transfer |T|/ε', value: c => c.st.maxTnl, format: expo(2), limits: [1e-9, 1e-6], roundoff: true },
  { label: 'energy budget residual (RK4-consistent)', value: c => c.st.maxEbal, format: expo(2), limits: [1e-5, 1e-3] },
  { label: 'enstrophy budget residual', value: c => c.st.maxZbal, format: expo(2), limits: [1e-4, 1e-2] },
  { label: 'resolution kmax·η', value: c => c.kmaxEta, format: fixed(2), limits: [1.0, 0.5], higher: true },
  { label: 'spectral tail E(kmax)/E(peak)', value: c => tailRatio(c.spec, c.kc), format: expo(2), limits: [1e-4, 1e-2] },
  { label: 'stretching: spectral vs physical', value: c => (c.o ? stretchGap(c.o, c.nu) : null), format: expo(2), limits: [1e-2, 1e-1] },
  { label: 'cutoff pile-up E(k)/E(0.8kmax)', value: c => c.pile, format: fixed(2), limits: [1.2, 2.0] },
  { label: 'max |ω|', text: c => c.d.omMax.toFixed(4) + (c.o && c.o.omMaxI != null ? ` (interpolated ${c.o.omMaxI.toFixed(4)})` : '') },
  { label: 'enstrophy Z', value: c => c.d.Z, format: fixed(5) },
  { label: 'stretching ⟨ω·S·ω⟩', value: c => c.d.Pspec, format: fixed(5) },
  { label: 'dissipation ε = 2νZ', value: c => c.d.eps, format: expo(4) },
  { label: '‖u‖_L³ (ESS criterion)', value: c => (c.o && c.o.uL3 != null ? c.o.uL3 : null), format: fixed(5) },
  { label: 'periodic-image gap in z', value: c => (c.o && c.o.imageGap != null ? c.o.imageGap : null), format: fixed(2) },
  { label: 'grid convergence', study: 'grid' },
  { label: 'time-step convergence', study: 'time' },
];

// The least healthy archived snapshot of each kind: `worse(a, b)` is true when b is less healthy than a (the first
// snapshot wins ties). `only` skips snapshots where the metric is undefined.
const INSTANT_CHECKS = [
  { label: 'worst instant: kmax·η', metric: (q, c) => q.kmaxEta, lower: true, format: fixed(2), limits: [1.0, 0.5], higher: true },
  { label: 'worst instant: spectral tail', metric: (q, c) => tailRatio(q.spectrum, c.kc), format: expo(2), limits: [1e-4, 1e-2] },
  { label: 'worst instant: stretching spectral vs physical', metric: (q, c) => stretchGap(q, c.nu), format: expo(2), limits: [1e-2, 1e-1] },
  { label: 'worst instant: cutoff pile-up', metric: q => q.pileUp, only: q => q.pileUp != null, format: fixed(2), limits: [1.2, 2.0] },
];

function endRow(chk, c) {
  if (chk.study) {
    const s = c.study && c.study[chk.study];
    return [chk.label, s ? s.text : 'not run', s ? s.verdict : 'N/A'];
  }
  if (chk.text) return [chk.label, chk.text(c), ''];
  const v = chk.value(c);
  if (chk.roundoff && c.fp32) return [chk.label + ROUNDOFF, chk.format(v), 'N/A'];
  if (v == null) return [chk.label, DASH, ''];
  return [chk.label, chk.format(v), chk.limits ? grade(v, chk.limits, chk.higher) : ''];
}

function instantRow(chk, snaps, c) {
  const pool = chk.only ? snaps.filter(chk.only) : snaps;
  if (!pool.length) return null;
  let w = pool[0], wv = chk.metric(w, c);
  for (const q of pool.slice(1)) {
    const v = chk.metric(q, c);
    if (chk.lower ? v < wv : v > wv) { w = q; wv = v; }
  }
  return [chk.label, `${chk.format(wv)} at t ${w.t.toFixed(2)}`, grade(wv, chk.limits, chk.higher)];
}

/**
 * health(study) of a solve (nslab.js:595-635). run: {N, kc, nu, dtFixed, st, last, spectrum, pileUp, fp32}; `spectrum`
 * is the last snapshot's or, without one, the current state's; `pileUp(spec)` the cutoff pile-up rule. An FP32 solve
 * (fp32 true) is graded as the GPU runner grades one (nslab_gpu.py:410-459) and its report carries
 * precision 'float32'.
 */
function healthReport(run, study) {
  const { N, kc, nu, st } = run;
  const d = run.last, o = st.outputs.length ? st.outputs[st.outputs.length - 1] : null;
  const spec = o ? o.spectrum : run.spectrum();
  const eta = Math.pow(nu * nu * nu / Math.max(d.eps, 1e-300), 0.25);
  const ctx = { N, kc, nu, dtFixed: run.dtFixed, st, d, o, spec, pile: run.pileUp(spec), study, fp32: !!run.fp32,
    kmaxEta: Math.sqrt(3) * kc * eta };
  const rows = END_CHECKS.map(chk => endRow(chk, ctx));
  const worstEnd = verdictOf(rows);
  if (st.outputs.length) {
    for (const chk of INSTANT_CHECKS) {
      const r = instantRow(chk, st.outputs, ctx);
      if (r) rows.push(r);
    }
  }
  if (ctx.fp32) return { rows, worst: verdictOf(rows), worstEnd, precision: 'float32', note: NOTE + NOTE_FP32 };
  return { rows, worst: verdictOf(rows), worstEnd, note: NOTE };
}

/** dossier(extra) (nslab.js:637-645): the record a run is audited from; `extra` adds or overrides keys. */
function dossierOf(run, instrument, extra) {
  const { st } = run;
  const record = {
    instrument,
    build: null,
    equations: EQUATIONS,
    parameters: { N: run.N, Re: run.Re, nu: run.nu, cfl: run.cfl, dt: run.dtFixed || 'adaptive', ic: st.ic, icParams: st.icParams, kc: run.kc },
    state: { t: st.t, steps: st.step },
    series: st.series,
    outputs: st.outputs,
    health: healthReport(run, extra && extra.study),
    disclaimer: DISCLAIMER,
  };
  return Object.assign(record, extra || {});
}

/**
 * Refinement verdict over a ladder of levels [{label, N, peak: {omMax}, epsPeak: {v, t}, Eend}] (nslab.js:652-666):
 * the relative changes of peak max|ω|, peak ε and E(t_end) between neighbouring levels; PASS when the last change
 * is under 1 % and smaller than the one before, WARN under 5 %, else FAIL; the observed order from the last two
 * changes and the refinement ratio (2 for a time-step ladder).
 */
function studySummary(levels, key) {
  if (levels.length < 2) return null;
  const change = (a, b) => Math.abs(a - b) / Math.max(Math.abs(b), 1e-300);
  const rows = levels.slice(1).map((b, i) => {
    const a = levels[i];
    return { from: a.label, to: b.label, omMax: change(a.peak.omMax, b.peak.omMax), eps: change(a.epsPeak.v, b.epsPeak.v),
      Eend: change(a.Eend, b.Eend), tPeak: Math.abs(a.epsPeak.t - b.epsPeak.t) };
  });
  const worst = rows.map(r => Math.max(r.omMax, r.eps, r.Eend));
  const now = worst[worst.length - 1], before = worst.length > 1 ? worst[worst.length - 2] : null;
  const refinement = () => (key === 'time' ? 2 : levels[levels.length - 1].N / levels[levels.length - 2].N);
  const order = before != null && now > 0 ? Math.log(before / now) / Math.log(refinement()) : null;
  const converging = before == null || now < before;
  const verdict = now < 0.01 && converging ? 'PASS' : now < 0.05 ? 'WARN' : 'FAIL';
  const r = rows[rows.length - 1], pct = (v, dp) => (v * 100).toFixed(dp);
  const text = `${levels.map(l => l.label).join(' → ')}: last change ω_max ${pct(r.omMax, 2)} %, ε_peak ${pct(r.eps, 2)} %, ` +
    `E(t_end) ${pct(r.Eend, 3)} %` + (order != null ? `, observed order ${order.toFixed(2)}` : '');
  return { rows, verdict, text, order };
}

module.exports = { healthReport, dossierOf, studySummary, grade, verdictOf, NOTE, NOTE_FP32, ROUNDOFF, EQUATIONS, DISCLAIMER };
