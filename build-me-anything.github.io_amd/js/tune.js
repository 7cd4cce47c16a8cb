'use strict';
/*
 * Autotuner of the kernel forms per grid size. For each N it times the compiled forms (nsl_tune_candidates) on this
 * device in two searches and writes the table the library compiles in (csrc/nsl_tune_table.inc) plus a JSON report:
 *   1. the x-pass form and its tile (row pairs per workgroup, persistent waves of the staged x pass) and whether
 *      the z-forward runs fused with the update, at the default pencil tile and radix schedules;
 *   2. from that pick, coordinate descent over the pencil passes' tile width and the radix schedules of the staged
 *      x pass and of the fused update: each one changed alone, the best change kept, twice.
 * Every form is checked to give the same step as the current form. A form replaces the current one only when it is
 * faster by more than the noise margin, so re-running on the same device keeps the table stable. Up to --pool-max
 * (256) each form is also timed in a pool of slots (the batched path: one launch per pass covers every slot), and a
 * form's score is the geometric mean of its single-solve and pool speeds relative to the reference form.
 *
 * The table is a build input: solvers then pick their forms by N alone (a solve gives the same bits alone and in
 * any pool). NSLAB_TUNE=<file> ("N xpass lp xwaves zfused pt xs zs" per line) overrides it at run time without a
 * rebuild.
 *
 *   node js/tune.js [--sizes 32,48,64,96,128,192,256,384,512] [--budget 0.6] [--margin 0.02] [--pool-max 256]
 *                   [--out csrc/nsl_tune_table.inc] [--report tune_report.json] [--dry]
 */
const fs = require('fs');
const path = require('path');
const addon = require(path.join(__dirname, '..', 'lib', 'nslab_napi.node'));

const FORM = { 1: 'row pairs', 2: 'rows', 3: 'wave', 4: 'staged wave', 5: 'plane pass' };
const REC = 21;   // doubles per step record of addon.run: t dt step ebal zbal last[8] d0[8]

function args() {
  const a = { sizes: [32, 48, 64, 96, 128, 192, 256, 384, 512], budget: 0.6, margin: 0.02, poolMax: 256, dry: false,
    out: path.join(__dirname, '..', 'csrc', 'nsl_tune_table.inc'), report: null };
  const v = process.argv.slice(2);
  for (let i = 0; i < v.length; i++) {
    const k = v[i];
    if (k === '--sizes') a.sizes = v[++i].split(',').map(Number);
    else if (k === '--budget') a.budget = +v[++i];
    else if (k === '--margin') a.margin = +v[++i];
    else if (k === '--out') a.out = v[++i];
    else if (k === '--report') a.report = v[++i];
    else if (k === '--pool-max') a.poolMax = +v[++i];
    else if (k === '--dry') a.dry = true;
    else throw new Error('unknown argument ' + k);
  }
  return a;
}

const now = () => Number(process.hrtime.bigint()) / 1e9;
const key = (c) => c.join(',');
const describe = (c) => FORM[c[0]] + (c[0] === 1 ? ' x' + c[1] : '') + (c[0] === 4 ? ' ' + (c[2] ? c[2] + '/CU' : 'one group per workgroup') : '') + (c[3] ? ', fused update' : '') +
  (c[4] ? ', alternative pencil tile' : '') + (c[5] === 1 ? ', alternative x-pass schedule' : c[5] === 2 ? ', one-wave x-pass groups' : '') + (c[6] ? ', alternative update schedule' : '');
const variants = (c) => (c[4] || 0) + (c[5] || 0) + (c[6] || 0);

// Steps/s of the forms `cfg` at N: TGV, nu = 1/1600, fixed dt; three untimed steps (their energy is the
// same-result check), then about `budget` seconds.
function timeForm(N, cfg, budget) {
  addon.tuneSet(N, ...cfg);
  const h = addon.create(N, 0, 1 / 1600, 0.4, 1e-3, 0, 0, 0, 0);
  try {
    const v = addon.run(h, 3, 0, 0);
    const E3 = v[2 * REC + 5];
    let n = 2, el = 0;
    for (;;) {
      const t0 = now();
      addon.run(h, n, 0, 0);
      el = now() - t0;
      if (el > 0.25 * budget || n >= 4096) break;
      n *= 2;
    }
    const reps = Math.max(1, Math.round(budget / Math.max(el, 1e-6)));
    const t0 = now();
    const r = addon.run(h, n * reps, 0, 0);
    const sec = now() - t0;
    return { sps: (n * reps) / sec, E: E3, steps: r.length / REC };
  } finally {
    addon.destroy(h);
  }
}

// Slots of the pool a form is timed in: about 16 GB of slots, 2 to 32.
const poolSlots = (N) => Math.max(2, Math.min(32, Math.floor(16e9 / addon.poolSlotBytes(N))));

// Solve-steps/s of the forms `cfg` at N in a pool of `slots` slots, the same solve as timeForm in every slot: 3 steps
// untimed (the step graph is captured on the first launch), then about `budget` seconds of 64-step chunks. The pool's
// device memory is freed before returning.
function timePool(N, cfg, slots, budget) {
  addon.tuneSet(N, ...cfg);
  const p = addon.poolCreate(N, slots);
  try {
    const ids = [];
    for (let i = 0; i < slots; i++) ids.push(addon.poolAdmit(p, N, 0, 1 / 1600, 0.4, 1e-3, 0, 0, 0, 0)[0]);
    const run = (n) => {
      for (const s of ids) addon.poolSetTarget(p, s, n, 0, 0);
      for (let left = n; left > 0; left -= 64) {
        const k = Math.min(64, left);
        addon.poolLaunch(p, k);
        addon.poolCollect(p, k);
      }
    };
    run(3);
    let n = 4, el = 0;
    for (;;) {
      const t0 = now();
      run(n);
      el = now() - t0;
      if (el > 0.25 * budget || n >= 1024) break;
      n *= 2;
    }
    const reps = Math.max(1, Math.round(budget / Math.max(el, 1e-6)));
    const t0 = now();
    run(n * reps);
    return (slots * n * reps) / (now() - t0);
  } finally {
    addon.poolDestroy(p);
  }
}

// Times the forms `order` (the reference first) alone and, up to --pool-max, in a pool; each checked against the
// reference's energy after 3 steps; score = speed relative to the reference.
function timeRows(N, order, a) {
  const rows = [];
  let ref = null;
  for (const c of order) {
    const m = timeForm(N, c, a.budget);
    if (ref === null) ref = m;
    const same = Math.abs(m.E - ref.E) / Math.abs(ref.E);
    const row = { form: c, what: describe(c), steps_per_s: m.sps, rel_diff_E: same };
    if (N <= a.poolMax) {
      row.pool_slots = poolSlots(N);
      row.pool_solve_steps_per_s = timePool(N, c, row.pool_slots, a.budget);
    }
    rows.push(row);
    if (!(same < 1e-11)) throw new Error(`N=${N} form ${describe(c)}: energy after 3 steps differs by ${same}`);
  }
  for (const r of rows) {
    r.score = r.steps_per_s / rows[0].steps_per_s;
    if (r.pool_solve_steps_per_s) r.score = Math.sqrt(r.score * r.pool_solve_steps_per_s / rows[0].pool_solve_steps_per_s);
  }
  return rows;
}

function tuneN(N, a) {
  const cur = addon.tuneGet(N);
  const cands = addon.tuneCandidates(N);
  // 1. x-pass forms and fusion at the default tile and schedules (the current form first: it is the reference)
  const base = cands.filter((c) => variants(c) === 0 && key(c) !== key(cur));
  const rows = timeRows(N, [cur].concat(base), a);
  let pick = choose(rows, a.margin);
  let gain = pick.score;
  // 2. tile width and schedules, one at a time from the pick (a variant must beat the pick by the margin)
  const variantRows = [];
  for (let round = 0; round < 2; round++) {
    const p = pick.form;
    const near = cands.filter((c) => key(c.slice(0, 4)) === key(p.slice(0, 4)) && key(c) !== key(p) &&
      [4, 5, 6].filter((i) => (c[i] || 0) !== (p[i] || 0)).length === 1 && !variantRows.some((r) => key(r.form) === key(c)));
    if (!near.length) break;
    const vr = timeRows(N, [p].concat(near), a);
    for (const r of vr.slice(1)) variantRows.push(Object.assign(r, { relative_to: describe(p) }));
    const best = vr.slice(1).reduce((b, r) => (r.score > b.score ? r : b), vr[0]);
    if (best === vr[0] || best.score < 1 + a.margin) break;
    gain *= best.score;
    pick = best;
  }
  addon.tuneSet(N);   // drop the override
  return { N, current: describe(cur), pick: pick.form, pick_what: pick.what, kept_current: key(pick.form) === key(cur),
    gain, candidates: rows, variants: variantRows };
}

// The form to keep, from timed rows (the current form first; `score` is the speed relative to it). Forms within the
// noise margin of the best score are a tie. A tie keeps the current form, except that the fused z-forward of the same x-pass form wins it: it moves
// 6 D complex values (D retained modes) fewer per rhs, which counts most in a pool, where every pass is
// bandwidth-bound. Without the current form among the tied ones, the fastest is taken.
function choose(rows, margin) {
  const best = rows.reduce((b, r) => (r.score > b.score ? r : b), rows[0]);
  const tied = (r) => r.score * (1 + margin) >= best.score;
  const cur = rows[0];
  if (!tied(cur)) return best;
  if (cur.form[3] === 0) {
    const twin = rows.find((r) => r.form[3] === 1 && key(r.form.slice(0, 3)) === key(cur.form.slice(0, 3)));
    if (twin && tied(twin)) return twin;
  }
  return cur;
}

// Rows of the sizes tuned now replace theirs; the rows of sizes not tuned in this run stay as the file has them.
function writeTable(file, results) {
  const rows = new Map();
  if (fs.existsSync(file)) {
    for (const line of fs.readFileSync(file, 'utf8').split('\n')) {
      const m = /^\{(\d+),/.exec(line);
      if (m) rows.set(+m[1], line);
    }
  }
  for (const r of results) rows.set(r.N, `{${r.N}, ${r.pick.concat([0, 0, 0]).slice(0, 7).join(', ')}},   // ${r.pick_what}`);
  const lines = ['// {N, xpass, lp, xwaves, zfused, pt, xs, zs}: written by js/tune.js (NSL_XP_PAIR 1, ROWS 2, WAVE 3, STAGED 4, PLANE 5).'];
  for (const N of [...rows.keys()].sort((a, b) => a - b)) lines.push(rows.get(N));
  fs.writeFileSync(file, lines.join('\n') + '\n');
}

if (require.main === module) {
  const a = args();
  const results = [];
  for (const N of a.sizes) {
    const r = tuneN(N, a);
    results.push(r);
    process.stdout.write(`N=${N}: ${r.pick_what}${r.kept_current ? ' (current kept)' : ''}, x${r.gain.toFixed(3)}; ` +
      r.candidates.concat(r.variants).map((c) => `${c.what} ${c.steps_per_s.toFixed(1)}` +
        (c.pool_solve_steps_per_s ? `/${c.pool_solve_steps_per_s.toFixed(0)} in ${c.pool_slots}` : '')).join(' | ') + '\n');
  }
  if (!a.dry) writeTable(a.out, results);
  if (a.report) fs.writeFileSync(a.report, JSON.stringify({ device: addon.version(), results }, null, 1));
}

module.exports = { tuneN, timeForm, choose, writeTable };
