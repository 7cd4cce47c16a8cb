/*
 * nslab_hip.h — C ABI of the MI355X (gfx950) NSLab solver core: libnslab_hip.so.
 *
 * Plain C types only (no HIP or torch types). Every entry point returns NSL_OK (0) or a negative status; the text of
 * the last failure on the calling thread is returned by nsl_last_error(). No exception crosses this boundary.
 * The Node host (nslab.js in the package) binds these through a hand-built N-API addon; Python tests bind them
 * through ctypes.
 *
 * Citations name the reference interface each entry point replaces (pocket-wind-tunnel/src/nslab.js = "nslab.js").
 * Spectral arrays at this boundary use the reference layout: six float64 planes (ux re, ux im, uy re, uy im, uz re,
 * uz im), each of NS = N*N*(N/2+1) values indexed (kz*N + ky)*(N/2+1) + kx (nslab.js:94, :184).
 */
#ifndef NSLAB_HIP_H
#define NSLAB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSL_OK 0
#define NSL_ERR_ARG -1      /* bad argument: grid size, unknown IC, null pointer (message as the reference's throw) */
#define NSL_ERR_HIP -2      /* HIP runtime failure (allocation, launch, copy) */
#define NSL_ERR_STATE -3    /* call not valid in the solver's current state */

/* Initial conditions (nslab.js:161-169 IC_INFO, setIC :250-292). */
#define NSL_IC_TGV 0
#define NSL_IC_TGV2D 1
#define NSL_IC_ABC 2
#define NSL_IC_TUBES 3
#define NSL_IC_RANDOM 4
/* Antiparallel tube pair along z (no reference counterpart: the memory-heavy benchmark's initial field). icp =
 * {Gamma, sigma, d, pert}: Gaussian cores of circulation +-Gamma and radius sigma, centrelines
 * (pi +- (d + 0.1 cos z), pi); velocity from the vorticity spectrally (u = i k x omega / k^2), then a seeded random
 * solenoidal perturbation (E(k) ~ k^4 exp(-2 (k/4)^2), seed 1) of rms velocity `pert` per component. */
#define NSL_IC_TUBESZ 5

typedef struct nsl_solver nsl_solver;

/* createSolver(opts) options (nslab.js:171-175). The caller applies the reference's defaults: nu = 1/(Re || 100),
 * cfl || 0.4, dt || 0 (0 = adaptive). icp: tubes {amp, sigma, sep, pert}; abc {A, B, C, -}; random {seed, k0, E0, -};
 * tgv/tgv2d ignore it. */
typedef struct {
  int32_t N;
  int32_t ic;
  double nu;
  double cfl;
  double dt;
  double icp[4];
} nsl_opts;

/* solver.last — diagnostics of the most recent rhs() input state (nslab.js:246). */
typedef struct {
  double E, Z, P, eps, Tnl, Pspec, omMax, uMax;
} nsl_rhsdiag;

/* One RK4 step as step() leaves it (nslab.js:312-333): st.t, st.dt, st.step after the step, the two budget
 * residuals, `last` (the rhs of the new state, the values record() appends) and d0 (the rhs of the state the step
 * started from, whose |Tnl|/eps feeds st.maxTnl). */
typedef struct {
  double t, dt;
  int64_t step;
  double ebal, zbal;
  nsl_rhsdiag last;
  nsl_rhsdiag d0;
} nsl_steprec;

/* createSolver(opts) (nslab.js:171): allocates the device state, builds the IC and runs the first rhs.
 * On success *out is a new solver; st.E0/st.Z0 are last.E/last.Z of the returned nsl_last(). */
int nsl_create(const nsl_opts *opts, nsl_solver **out);
/* Solver(N, Re, ic, icParams, cfl, dt, xp, fp32=True) of the GPU runner (nslab_gpu.py:56-65): the same solve with
 * every stepped array (state, RK4 registers, each pass's output) held in complex64, half the bytes of nsl_create's
 * FP64 solve; the transforms and the product u x omega run in single precision as the runner's do, the update
 * (projection, viscous term, RK4) and the sums in FP64, so results carry FP32 rounding in every pass. For N = 32,
 * 48, 64, 96, 128, 192, 256, 384, 512, 1024 (NSL_ERR_ARG otherwise). Never a pool slot. The host views (initial
 * field, nsl_get_state/nsl_set_state, nsl_norms and the diagnostics) allocate an FP64 copy of the state and its
 * transform scratch for the duration of the call, and nsl_spectrum_noalloc refuses an FP32 solve with
 * NSL_ERR_STATE. */
int nsl_create_fp32(const nsl_opts *opts, nsl_solver **out);
/* 64 or 32: the storage precision of a solve (the runner's fp32 flag, nslab_gpu.py:59). */
int nsl_precision(const nsl_solver *s);
/* solver.setIC(kind, prm) (nslab.js:250): resets t and step to 0 and recomputes last. */
int nsl_set_ic(nsl_solver *s, int32_t ic, const double icp[4]);
void nsl_destroy(nsl_solver *s);

/* run(nSteps, tEnd) (nslab.js:334): up to n steps, stopping before a step once t >= tEnd - 1e-12 (has_tend != 0).
 * recs (may be NULL) receives one nsl_steprec per step taken; *n_done the count. */
int nsl_run(nsl_solver *s, int64_t n, int32_t has_tend, double tend, nsl_steprec *recs, int64_t *n_done);

/* Clock the step loop reads: st.t, st.step and st.E0 (the energy the budget residual is normalised by). Callers
 * that assign st (checkpoint restore, nslab-worker) push the new values here before stepping. */
int nsl_set_clock(nsl_solver *s, double t, int64_t step, double E0);
int nsl_get_clock(nsl_solver *s, double *t, int64_t *step, double *dt);

/* solver.last (nslab.js:648 getter). */
int nsl_last(nsl_solver *s, nsl_rhsdiag *out);
/* solver.norms() (nslab.js:214-222): out = {E, Z, P}. */
int nsl_norms(nsl_solver *s, double out[3]);

/* solver._S (nslab.js:184, 648): copy the state out to / in from six host planes of NS float64 each. Modes outside
 * the 2/3 mask are identically zero in every state the solver produces (nslab.js:235, :289) and are not stored;
 * set_state drops them. set_state re-evaluates rhs so `last` describes the new state, as step() would at entry. */
int nsl_get_state(nsl_solver *s, double *const planes[6]);
int nsl_set_state(nsl_solver *s, const double *const planes[6]);

/* solver.diagnose() (nslab.js:339-392). out receives NSL_DIAG_LEN + nb doubles (nb = nsl_spectrum_len):
 *  0 t, 1 step, 2 E, 3 Z, 4 P, 5 eps, 6 omMax, 7 uMax, 8 Pspec, 9 Pphys, 10 Tnl, 11 skew, 12 divMax, 13 eta,
 *  14 kmaxEta, 15 Rel, 16 lambda, 17 skewIso, 18..20 align, 21 uL3, 22 pileUp (NaN = null), 23 zCentroid,
 *  24 zExtent, 25 imageGap, 26 zBands, 27 omMaxI (NaN = not run), 28..30 omMaxIpos, 31 omMaxNode,
 *  32..79 alignHist[3][16], 80.. spectrum[nb].
 * do_interp != 0 runs interpMax when N <= 128 (nslab.js:389-390). Updates the solver's divMax (returned by
 * nsl_get_divmax). */
#define NSL_DIAG_LEN 80
int nsl_diagnose(nsl_solver *s, int32_t do_interp, double *out);
int nsl_get_divmax(nsl_solver *s, double *divmax);
int nsl_set_divmax(nsl_solver *s, double divmax);
/* solver.spectrum() (nslab.js:414-420) and its length ceil(sqrt(3) kc) + 2. */
int nsl_spectrum_len(nsl_solver *s);
int nsl_spectrum(nsl_solver *s, double *Ek);
/* nsl_spectrum without any allocation. No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that forbid hipMalloc). Same kernels and launch
 * shape as nsl_spectrum, so the same bits; its per-workgroup partials live in a fixed 4 MiB device array per GPU
 * (held until the process ends), which covers every grid up to 768^3; larger grids return NSL_ERR_ARG. Calls on one
 * GPU share that array: calls that overlap in time on one GPU (other threads, other solvers) race. */
int nsl_spectrum_noalloc(nsl_solver *s, double *Ek);
/* solver.interpMax(top, iters) (nslab.js:421-536): out = {omMaxI, x, y, z, omMaxNode}. Needs the vorticity field of
 * a preceding nsl_diagnose. */
int nsl_interp_max(nsl_solver *s, int32_t top, int32_t iters, double out[5]);
/* solver.fields / slice() source (nslab.js:431, 566-583): physical field `which` (0..2 u v w, 3..5 omega after
 * diagnose or u x omega after a step, 6 Q, 7 stretching) as N^3 float64 in (z*N + y)*N + x order. */
int nsl_field(nsl_solver *s, int32_t which, double *out);
/* solver.exactError() (nslab.js:585-594): returns 1 and fills {linf, rel} for abc/tgv2d, 0 for other ICs. */
int nsl_exact_error(nsl_solver *s, double out[2]);

/* solver.slice(kind, axis, index) (nslab.js:566-583): an N x N float32 plane of a physical quantity, out[b*N + a],
 * computed on the GPU. axis 2 is the plane z = index (a = x, b = y), 1 is y = index (a = x, b = z), anything else
 * x = index (a = y, b = z); index is clamped to 0..N-1. |omega|, |u|, u.u/2, u, v, w and omega are those of the
 * current state; Q and the stretching density those of the last nsl_diagnose (zero before one). Other kinds give a
 * plane of zeros, as the reference's slice does for an unknown kind. */
#define NSL_SLICE_VORT 0
#define NSL_SLICE_SPEED 1
#define NSL_SLICE_KE 2
#define NSL_SLICE_Q 3
#define NSL_SLICE_STRETCH 4
#define NSL_SLICE_U 5
#define NSL_SLICE_V 6
#define NSL_SLICE_W 7
#define NSL_SLICE_OMX 8
#define NSL_SLICE_OMY 9
#define NSL_SLICE_OMZ 10
int nsl_slice(nsl_solver *s, int32_t kind, int32_t axis, int32_t index, float *out);
/* solver.imageDiag() (nslab.js:537-550) of the current state's vorticity: out = {zCentroid, zExtent, imageGap,
 * zBands}. */
int nsl_image_diag(nsl_solver *s, double out[4]);
/* The cutoff pile-up of a shell spectrum (cutoffPileUp, nslab.js:557-564) on a grid with kc = floor(N/3): the largest
 * E(k)/E(k8) for k8 = round(0.8 kc) <= k <= kc. Returns 1 and *out, or 0 (the reference's null) when E(k8) is not
 * above 100 eps^2 times the largest E(k), k >= 1; NSL_ERR_ARG unless 1 <= kc < nb. Host arithmetic only. */
int nsl_cutoff_pile_up(int32_t kc, const double *spec, int32_t nb, double *out);
/* The same guard in a working precision of machine epsilon eps (the GPU runner's pileUp with specFloor = 100 eps^2,
 * nslab_gpu.py:61-65,357-372): eps = 1.1920929e-07 for an FP32 solve. NSL_ERR_ARG unless 0 < eps < 1. */
int nsl_cutoff_pile_up_eps(int32_t kc, const double *spec, int32_t nb, double eps, double *out);
/* solver.N: the grid size of a solver or pool slot. */
int nsl_grid_size(const nsl_solver *s);

/* Bytes of device memory the solver holds (admission accounting). */
int64_t nsl_device_bytes(const nsl_solver *s);

/* Per-kernel timing (no reference counterpart: the measurement hook the benchmark uses). With profiling on, every
 * launch of a step is bracketed by HIP events on the solver's stream; nsl_profile_read returns the summed
 * milliseconds and launch counts per kernel class since nsl_profile(s, 1) reset them. */
#define NSL_PROF_ZINV 0      /* P1: z-inverse + curl */
#define NSL_PROF_YINV 1      /* P2: y-inverse */
#define NSL_PROF_XNL 2       /* P3: x C2R, u x omega, x R2C */
#define NSL_PROF_YFWD 3      /* P4: y-forward */
#define NSL_PROF_ZFWD 4      /* P5: z-forward */
#define NSL_PROF_UPDATE1 5   /* projection + RK4 stage 1 */
#define NSL_PROF_UPDATE23 6  /* projection + RK4 stages 2, 3 */
#define NSL_PROF_UPDATE4 7   /* projection + RK4 stage 4 + norms */
#define NSL_PROF_FINALIZE 8  /* fixed-order sum of the partials, step bookkeeping */
#define NSL_PROF_CLASSES 9
int nsl_profile(nsl_solver *s, int32_t on);
int nsl_profile_read(nsl_solver *s, double *ms, int64_t *count);

/* Slot pools: many solves of one N resident on the device and stepped together. The reference runs the levels of
 * a study one createSolver/step loop at a time (nslab-worker.js:42-49, startLevel/loop) and the GPU runner one
 * solve per process; a pool replaces those loops when several solves are in flight. Every pass is one launch over
 * all slots (gridDim.z = slots); a slot's arithmetic is the single solve's, so its results do not depend on what
 * the other slots hold. One RK4 step of the pool is a captured hipGraph.
 *   nsl_pool_slot_bytes(N)         device bytes one slot of grid N occupies (admission by HBM footprint)
 *   nsl_pool_create(N, slots, &p)  allocates `slots` slots, all free
 *   nsl_pool_admit(p, opts, &i)    createSolver(opts) into a free slot (opts.N must equal the pool's N)
 *   nsl_pool_slot(p, i, &s)        the slot as a solver handle: every nsl_* solver call above works on it
 *                                  (state, clock, last, diagnose, spectrum, nsl_run); nsl_destroy ignores it
 *   nsl_pool_set_target(p, i, n, has_tend, tend)   the slot runs n more steps (stopping at tend), as run(n, tEnd)
 *   nsl_pool_launch(p, k)          enqueues k <= 64 steps of every slot with steps left; returns at once
 *   nsl_pool_ready(p)              1 once the launched steps have finished (collect will not wait), else 0
 *   nsl_pool_collect(p, recs, cap, n_done, &running)   waits; per slot i: n_done[i] steps and their records at
 *                                  recs[i*cap ..] (recs may be null); running = slots with steps left
 *   nsl_pool_release(p, i)         frees the slot
 * Between nsl_pool_launch and nsl_pool_collect a slot that took part in the launched chunk belongs to it:
 * nsl_pool_set_target, nsl_run, nsl_set_state and nsl_set_ic on it return NSL_ERR_STATE. Slots admitted or
 * re-targeted after the launch were not part of it and may be set up while it runs. */
typedef struct nsl_pool nsl_pool;
int64_t nsl_pool_slot_bytes(int32_t N);
/* HBM copy rate of the current GPU (no reference counterpart: the practical ceiling the benchmark reports beside the
 * 8 TB/s peak): `reps` copies of a `bytes` buffer by a plain copy kernel, *gbs = bytes read + written per second / 1e9.
 * Allocates and frees two buffers of `bytes`. */
int nsl_copy_probe(int64_t bytes, int32_t reps, double *gbs);
/* Free and total device memory of the current GPU (the admission budget of a request loop). */
int nsl_device_memory(int64_t *free_bytes, int64_t *total_bytes);
int nsl_pool_create(int32_t N, int32_t slots, nsl_pool **out);
void nsl_pool_destroy(nsl_pool *p);
int nsl_pool_info(nsl_pool *p, int32_t *N, int32_t *slots, int32_t *used, int32_t *graphed, int64_t *bytes);
int nsl_pool_admit(nsl_pool *p, const nsl_opts *opts, int32_t *slot);
int nsl_pool_slot(nsl_pool *p, int32_t slot, nsl_solver **view);
int nsl_pool_set_target(nsl_pool *p, int32_t slot, int64_t n, int32_t has_tend, double tend);
int nsl_pool_launch(nsl_pool *p, int32_t steps);
int nsl_pool_ready(nsl_pool *p);
int nsl_pool_collect(nsl_pool *p, nsl_steprec *recs, int32_t recs_per_slot, int64_t *n_done, int32_t *n_running);
int nsl_pool_release(nsl_pool *p, int32_t slot);

/* Pool groups: pools of different N stepped as one. The reference steps one solve at a time (nslab-worker.js:42-49);
 * an ensemble of uneven grids is here one pool per N, and a group makes one RK4 step of all of them one captured
 * hipGraph (a fork/join capture of every pool's step), so an ensemble step is a single graph launch.
 *   nsl_group_create(pools, n, &g)  captures the graph (no pool may have a chunk in flight); the pools stay usable
 *                                   alone, and must outlive the group
 *   nsl_group_launch(g, k)          k <= 64 steps of every slot with steps left in every pool of the group; returns
 *                                   at once; then nsl_pool_ready / nsl_pool_collect on each pool as after
 *                                   nsl_pool_launch (each pool's stream waits for the group's work)
 *   nsl_group_destroy(g)
 * Slot results are bit-identical to the same pools launched one by one. */
typedef struct nsl_group nsl_group;
int nsl_group_create(nsl_pool *const *pools, int32_t n, nsl_group **out);
int nsl_group_launch(nsl_group *g, int32_t steps);
void nsl_group_destroy(nsl_group *g);

/* NS.fft1 / NS.Spectral (nslab.js:46-155) on the GPU: `lines` independent length-N lines (split re/im, line-major)
 * transformed in place, sign -1 forward, +1 inverse, unnormalised. */
int nsl_fft1(int32_t N, int32_t sign, int64_t lines, double *re, double *im);
/* Spectral.forward: real f[N^3] -> re, im [NS], normalised by 1/N^3 so f = sum u_k e^{ik.x}. */
int nsl_spectral_forward(int32_t N, const double *f, double *re, double *im);
/* Spectral.inverse: re, im [NS] -> real f[N^3], unnormalised. */
int nsl_spectral_inverse(int32_t N, const double *re, const double *im, double *f);

/* factorize(N) (nslab.js:28-35): radices into f[0..return), or NSL_ERR_ARG with the reference's message. */
int nsl_factorize(int32_t N, int32_t *f, int32_t cap);

/* ---- Kernel forms per grid size (autotuner). No counterpart in the reference: it has one code path. Tile widths
 * and radix schedules are picked per N by the project's autotuner (js/tune.js). A solver's forms depend on N only
 * and are fixed when it (or its pool) is created, so a solve gives the same bits alone and in any pool.
 *   xpass : x-pass form, 1 row pairs (16x16-style schedules, lp row pairs per workgroup), 2 one row per workgroup,
 *           3 one wave per row group (8x8x8 at 512, 8x4x8 at 256), 4 the same with its input staged through LDS,
 *           5 the plane pass: y-inverse, x pass and y-forward in one workgroup per z-plane (32, 48, 64)
 *   lp    : row pairs per workgroup of form 1 (1, 2, 4, 8)
 *   xwaves: persistent waves per CU of form 4 (0: one row group per workgroup)
 *   zfused: 1 = z-forward fused with the spectral update, 0 = two kernels
 *   ptile : pencil-pass tile width, 0 the default lines per workgroup of N, 1 the alternative (32, 48, 64: 8 lines;
 *           96: 40; 128, 192: 24; no alternative at other sizes)
 *   xsched: staged x pass (xpass 4 only): 0 the default radix schedule, 1 the alternative, 2 the default schedule in
 *           one-wave row groups with idle lanes in the first and last stages (48, 96, 192, 384, whose default groups
 *           are three waves)
 *   zsched: radix schedule of the fused z-forward, 0 the default, 1 the alternative (zfused 1 only) */
typedef struct {
  int32_t xpass, lp, xwaves, zfused;
  int32_t ptile, xsched, zsched;
} nsl_tune;
/* The forms solvers of this N use: an nsl_tune_set override, else the NSLAB_TUNE file, else the tuned table. */
int nsl_tune_get(int32_t N, nsl_tune *out);
/* Override for solvers and pools created afterwards (cfg == NULL drops it); NSL_ERR_ARG if not compiled for N. */
int nsl_tune_set(int32_t N, const nsl_tune *cfg);
/* Every compiled form of N into out[0..min(return, cap)); returns the count. */
int nsl_tune_candidates(int32_t N, nsl_tune *out, int32_t cap);
/* The x pass's launch plan for form cfg at grid N on a device of `cus` compute units (0: the current device's):
 * *workgroups receives how many workgroups it launches, each leaving one maxima pair. NSL_ERR_STATE when the form
 * is not compiled for N or when those pairs exceed cap_pairs. A solver checks this plan against its buffer of N^2
 * pairs before it enqueues the pass, so an oversized grid is refused with nothing launched. Host arithmetic only. */
int nsl_xpass_plan(int32_t N, const nsl_tune *cfg, int32_t cus, int64_t cap_pairs, int64_t *workgroups);

/* Text of the last error on this thread ("" if none). */
const char *nsl_last_error(void);
/* Version string of the library. */
const char *nsl_version(void);

#ifdef __cplusplus
}
#endif
#endif
