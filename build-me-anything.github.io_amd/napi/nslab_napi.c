/*
 * nslab_napi.node — N-API binding of the C ABI in include/nslab_hip.h for the Node host module js/nslab.js.
 * Built with gcc against node_api.h only (no node-gyp); links libnslab_hip.so from the same directory.
 * Every nonzero status from the library becomes a JS Error carrying nsl_last_error()'s text.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "nslab_hip.h"

#define CALL(env, expr)                                                     \
  do {                                                                      \
    if ((expr) != napi_ok) {                                                \
      napi_throw_error((env), NULL, "N-API call failed: " #expr);           \
      return NULL;                                                          \
    }                                                                       \
  } while (0)

#define ARGS(n)                                                             \
  size_t argc = (n);                                                        \
  napi_value argv[(n) > 0 ? (n) : 1];                                       \
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL))

static napi_value throw_status(napi_env env, int rc) {
  const char *m = nsl_last_error();
  napi_throw_error(env, NULL, (m && m[0]) ? m : "nslab-hip error");
  (void)rc;
  return NULL;
}

static int get_f64(napi_env env, napi_value v, double *out) { return napi_get_value_double(env, v, out) == napi_ok; }
static int get_i64(napi_env env, napi_value v, int64_t *out) { return napi_get_value_int64(env, v, out) == napi_ok; }

/* Float64Array -> pointer and element count */
static int get_f64arr(napi_env env, napi_value v, double **p, size_t *n) {
  napi_typedarray_type t;
  size_t len, off;
  void *data;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off) != napi_ok || t != napi_float64_array) return 0;
  *p = (double *)data;
  *n = len;
  return 1;
}

static napi_value new_f64arr(napi_env env, size_t n, double **data) {
  napi_value ab, ta;
  void *p;
  if (napi_create_arraybuffer(env, n * sizeof(double), &p, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, napi_float64_array, n, ab, 0, &ta) != napi_ok) return NULL;
  *data = (double *)p;
  return ta;
}

/* A JS solver handle: the solver, and whether the handle owns it (a pool slot view does not). destroy(h) frees an
 * owned solver at once (its device memory with it); the finaliser frees whatever is left when h is collected. */
typedef struct {
  nsl_solver *s;
  int owned;
} solver_ref;

static void finalize_solver(napi_env env, void *data, void *hint) {
  (void)env; (void)hint;
  solver_ref *r = (solver_ref *)data;
  if (r->owned && r->s) nsl_destroy(r->s);
  free(r);
}

static napi_value wrap_solver(napi_env env, nsl_solver *s, int owned) {
  solver_ref *r = (solver_ref *)malloc(sizeof *r);
  if (!r) { if (owned) nsl_destroy(s); napi_throw_error(env, NULL, "out of host memory"); return NULL; }
  r->s = s;
  r->owned = owned;
  napi_value ext;
  if (napi_create_external(env, r, finalize_solver, NULL, &ext) != napi_ok) {
    finalize_solver(env, r, NULL);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_external");
    return NULL;
  }
  return ext;
}

static solver_ref *get_ref(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return (solver_ref *)p;
}

static nsl_solver *get_solver(napi_env env, napi_value v) {
  solver_ref *r = get_ref(env, v);
  return r ? r->s : NULL;
}

#define SOLVER(idx)                                                         \
  nsl_solver *s = get_solver(env, argv[idx]);                               \
  if (!s) { napi_throw_type_error(env, NULL, "not a solver handle"); return NULL; }

/* create(N, ic, nu, cfl, dt, icp0..icp3) -> handle */
/* create(N, ic, nu, cfl, dt, p0, p1, p2, p3[, fp32]): fp32 truthy selects complex64 storage (nsl_create_fp32) */
static napi_value js_create(napi_env env, napi_callback_info info) {
  ARGS(10);
  nsl_opts o;
  memset(&o, 0, sizeof o);
  int64_t N = 0, ic = 0, fp32 = 0;
  if (argc < 9 || !get_i64(env, argv[0], &N) || !get_i64(env, argv[1], &ic) || !get_f64(env, argv[2], &o.nu) ||
      !get_f64(env, argv[3], &o.cfl) || !get_f64(env, argv[4], &o.dt)) {
    napi_throw_type_error(env, NULL, "create(N, ic, nu, cfl, dt, p0, p1, p2, p3[, fp32])");
    return NULL;
  }
  for (int i = 0; i < 4; i++) if (!get_f64(env, argv[5 + i], &o.icp[i])) o.icp[i] = 0;
  if (argc >= 10 && !get_i64(env, argv[9], &fp32)) fp32 = 0;
  o.N = (int32_t)N;
  o.ic = (int32_t)ic;
  nsl_solver *s = NULL;
  int rc = fp32 ? nsl_create_fp32(&o, &s) : nsl_create(&o, &s);
  if (rc < 0) return throw_status(env, rc);
  return wrap_solver(env, s, 1);
}

/* destroy(h): frees the solver now (a pool slot view is left to its pool); later calls on h throw */
static napi_value js_destroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  solver_ref *r = get_ref(env, argv[0]);
  if (!r) { napi_throw_type_error(env, NULL, "not a solver handle"); return NULL; }
  if (r->owned && r->s) nsl_destroy(r->s);
  r->s = NULL;
  return NULL;
}

static napi_value js_set_ic(napi_env env, napi_callback_info info) {
  ARGS(6);
  SOLVER(0);
  int64_t ic = 0;
  double p[4] = {0, 0, 0, 0};
  get_i64(env, argv[1], &ic);
  for (int i = 0; i < 4; i++) get_f64(env, argv[2 + i], &p[i]);
  int rc = nsl_set_ic(s, (int32_t)ic, p);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

/* run(h, n, hasTEnd, tEnd) -> Float64Array(nDone * 21): t dt step ebal zbal last[8] d0[8] */
static napi_value js_run(napi_env env, napi_callback_info info) {
  ARGS(4);
  SOLVER(0);
  int64_t n = 0, has = 0;
  double tend = 0;
  get_i64(env, argv[1], &n);
  get_i64(env, argv[2], &has);
  get_f64(env, argv[3], &tend);
  if (n < 0) n = 0;
  nsl_steprec *recs = (nsl_steprec *)calloc(n > 0 ? (size_t)n : 1, sizeof(nsl_steprec));
  if (!recs) { napi_throw_error(env, NULL, "out of host memory"); return NULL; }
  int64_t done = 0;
  int rc = nsl_run(s, n, (int32_t)has, tend, recs, &done);
  if (rc < 0) { free(recs); return throw_status(env, rc); }
  double *d;
  napi_value out = new_f64arr(env, (size_t)done * 21, &d);
  for (int64_t i = 0; i < done; i++) {
    const nsl_steprec *r = &recs[i];
    double *o = d + i * 21;
    o[0] = r->t; o[1] = r->dt; o[2] = (double)r->step; o[3] = r->ebal; o[4] = r->zbal;
    memcpy(o + 5, &r->last, 8 * sizeof(double));
    memcpy(o + 13, &r->d0, 8 * sizeof(double));
  }
  free(recs);
  return out;
}

static napi_value js_last(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  nsl_rhsdiag d;
  int rc = nsl_last(s, &d);
  if (rc < 0) return throw_status(env, rc);
  double *p;
  napi_value out = new_f64arr(env, 8, &p);
  memcpy(p, &d, sizeof d);
  return out;
}

static napi_value js_set_clock(napi_env env, napi_callback_info info) {
  ARGS(4);
  SOLVER(0);
  double t = 0, E0 = 0;
  int64_t step = 0;
  get_f64(env, argv[1], &t);
  get_i64(env, argv[2], &step);
  get_f64(env, argv[3], &E0);
  int rc = nsl_set_clock(s, t, step, E0);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

static napi_value js_get_clock(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  double t, dt;
  int64_t step;
  int rc = nsl_get_clock(s, &t, &step, &dt);
  if (rc < 0) return throw_status(env, rc);
  double *p;
  napi_value out = new_f64arr(env, 3, &p);
  p[0] = t; p[1] = (double)step; p[2] = dt;
  return out;
}

static napi_value js_norms(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  double *p;
  napi_value out = new_f64arr(env, 3, &p);
  int rc = nsl_norms(s, p);
  if (rc < 0) return throw_status(env, rc);
  return out;
}

/* getState(h, [6 Float64Array]) / setState(h, [6 Float64Array]) */
static int planes_of(napi_env env, napi_value arr, double *planes[6], size_t want) {
  for (uint32_t i = 0; i < 6; i++) {
    napi_value e;
    size_t n;
    if (napi_get_element(env, arr, i, &e) != napi_ok || !get_f64arr(env, e, &planes[i], &n) || n < want) return 0;
  }
  return 1;
}

/* Complex modes per half-spectrum plane of the solver's grid, N * N * (N/2 + 1). */
static size_t plane_len(nsl_solver *s) {
  const size_t N = (size_t)nsl_grid_size(s);
  return N * N * (N / 2 + 1);
}

static napi_value js_get_state(napi_env env, napi_callback_info info) {
  ARGS(3);
  SOLVER(0);
  const size_t NS = plane_len(s);
  double *pl[6];
  if (!planes_of(env, argv[1], pl, NS)) { napi_throw_range_error(env, NULL, "state needs 6 Float64Array of N*N*(N/2+1)"); return NULL; }
  int rc = nsl_get_state(s, pl);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

static napi_value js_set_state(napi_env env, napi_callback_info info) {
  ARGS(3);
  SOLVER(0);
  const size_t NS = plane_len(s);
  double *pl[6];
  if (!planes_of(env, argv[1], pl, NS)) { napi_throw_range_error(env, NULL, "state needs 6 Float64Array of N*N*(N/2+1)"); return NULL; }
  int rc = nsl_set_state(s, (const double *const *)pl);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

static napi_value js_diagnose(napi_env env, napi_callback_info info) {
  ARGS(2);
  SOLVER(0);
  int64_t interp = 1;
  get_i64(env, argv[1], &interp);
  int nb = nsl_spectrum_len(s);
  if (nb < 0) return throw_status(env, nb);
  double *p;
  napi_value out = new_f64arr(env, NSL_DIAG_LEN + (size_t)nb, &p);
  int rc = nsl_diagnose(s, (int32_t)interp, p);
  if (rc < 0) return throw_status(env, rc);
  return out;
}

static napi_value js_spectrum(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  int nb = nsl_spectrum_len(s);
  if (nb < 0) return throw_status(env, nb);
  double *p;
  napi_value out = new_f64arr(env, (size_t)nb, &p);
  int rc = nsl_spectrum(s, p);
  if (rc < 0) return throw_status(env, rc);
  return out;
}

static napi_value js_interp_max(napi_env env, napi_callback_info info) {
  ARGS(3);
  SOLVER(0);
  int64_t top = 24, iters = 12;
  get_i64(env, argv[1], &top);
  get_i64(env, argv[2], &iters);
  double *p;
  napi_value out = new_f64arr(env, 5, &p);
  int rc = nsl_interp_max(s, (int32_t)top, (int32_t)iters, p);
  if (rc < 0) return throw_status(env, rc);
  return out;
}

/* field(h, which, Float64Array(N^3)) */
static napi_value js_field(napi_env env, napi_callback_info info) {
  ARGS(3);
  SOLVER(0);
  int64_t which = 0;
  get_i64(env, argv[1], &which);
  double *p;
  size_t n;
  if (!get_f64arr(env, argv[2], &p, &n)) { napi_throw_type_error(env, NULL, "field needs a Float64Array"); return NULL; }
  const size_t N = (size_t)nsl_grid_size(s);
  if (n < N * N * N) { napi_throw_range_error(env, NULL, "field needs a Float64Array of N^3"); return NULL; }
  int rc = nsl_field(s, (int32_t)which, p);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

static napi_value js_exact_error(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  double e[2];
  int rc = nsl_exact_error(s, e);
  if (rc < 0) return throw_status(env, rc);
  if (rc == 0) { napi_value nul; CALL(env, napi_get_null(env, &nul)); return nul; }
  double *p;
  napi_value out = new_f64arr(env, 2, &p);
  p[0] = e[0]; p[1] = e[1];
  return out;
}

/* slice(h, kind, axis, index) -> Float32Array(N*N) computed on the GPU (kind, axis as nsl_slice) */
static napi_value js_slice(napi_env env, napi_callback_info info) {
  ARGS(4);
  SOLVER(0);
  int64_t kind = -1, axis = 0, index = 0;
  get_i64(env, argv[1], &kind);
  get_i64(env, argv[2], &axis);
  get_i64(env, argv[3], &index);
  const size_t N = (size_t)nsl_grid_size(s);
  napi_value ab, ta;
  void *p;
  CALL(env, napi_create_arraybuffer(env, N * N * sizeof(float), &p, &ab));
  CALL(env, napi_create_typedarray(env, napi_float32_array, N * N, ab, 0, &ta));
  int rc = nsl_slice(s, (int32_t)kind, (int32_t)axis, (int32_t)index, (float *)p);
  if (rc < 0) return throw_status(env, rc);
  return ta;
}

/* imageDiag(h) -> [zCentroid, zExtent, imageGap, zBands] */
static napi_value js_image_diag(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  double *p;
  napi_value out = new_f64arr(env, 4, &p);
  int rc = nsl_image_diag(s, p);
  if (rc < 0) return throw_status(env, rc);
  return out;
}

/* cutoffPileUp(kc, Float64Array spectrum) -> number, or null below the noise floor */
/* cutoffPileUp(kc, spectrum[, eps]): eps = machine epsilon of the working precision (FP64 when absent) */
static napi_value js_cutoff_pile_up(napi_env env, napi_callback_info info) {
  ARGS(3);
  int64_t kc = 0;
  double *sp;
  size_t nb;
  if (!get_i64(env, argv[0], &kc) || !get_f64arr(env, argv[1], &sp, &nb)) {
    napi_throw_type_error(env, NULL, "cutoffPileUp(kc, Float64Array spectrum[, eps])");
    return NULL;
  }
  double v = 0, eps = 0;
  const int has_eps = argc >= 3 && get_f64(env, argv[2], &eps);
  int rc = has_eps ? nsl_cutoff_pile_up_eps((int32_t)kc, sp, (int32_t)nb, eps, &v)
                   : nsl_cutoff_pile_up((int32_t)kc, sp, (int32_t)nb, &v);
  if (rc < 0) return throw_status(env, rc);
  napi_value out;
  if (rc == 0) CALL(env, napi_get_null(env, &out));
  else CALL(env, napi_create_double(env, v, &out));
  return out;
}

/* precision(h): 64 or 32 */
static napi_value js_precision(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  const int rc = nsl_precision(s);
  if (rc < 0) return throw_status(env, rc);
  napi_value v;
  CALL(env, napi_create_int32(env, rc, &v));
  return v;
}

static napi_value js_device_bytes(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  napi_value v;
  CALL(env, napi_create_double(env, (double)nsl_device_bytes(s), &v));
  return v;
}

/* fft1(N, sign, re, im): in place on Float64Arrays holding lines of N */
static napi_value js_fft1(napi_env env, napi_callback_info info) {
  ARGS(4);
  int64_t N = 0, sign = 0;
  double *re, *im;
  size_t nr, ni;
  get_i64(env, argv[0], &N);
  get_i64(env, argv[1], &sign);
  if (!get_f64arr(env, argv[2], &re, &nr) || !get_f64arr(env, argv[3], &im, &ni) || nr != ni || N <= 0 || nr % (size_t)N) {
    napi_throw_type_error(env, NULL, "fft1(N, sign, Float64Array re, Float64Array im)");
    return NULL;
  }
  int rc = nsl_fft1((int32_t)N, (int32_t)sign, (int64_t)(nr / (size_t)N), re, im);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

/* The library reads or writes N^3 reals and N*N*(N/2+1) modes whatever the arrays hold: refuse shorter ones (a grid
 * the library rejects is left to its own error). */
static int spectral_sizes_ok(int64_t N, size_t nreal, size_t nre, size_t nim) {
  if (N < 4 || N > 4096) return 1;
  const size_t n = (size_t)N, ns = n * n * (n / 2 + 1);
  return nreal >= n * n * n && nre >= ns && nim >= ns;
}

static napi_value js_spectral_forward(napi_env env, napi_callback_info info) {
  ARGS(4);
  int64_t N = 0;
  double *f, *re, *im;
  size_t n1, n2, n3;
  get_i64(env, argv[0], &N);
  if (!get_f64arr(env, argv[1], &f, &n1) || !get_f64arr(env, argv[2], &re, &n2) || !get_f64arr(env, argv[3], &im, &n3)) {
    napi_throw_type_error(env, NULL, "spectralForward(N, f, re, im)");
    return NULL;
  }
  if (!spectral_sizes_ok(N, n1, n2, n3)) {
    napi_throw_range_error(env, NULL, "spectralForward needs f of N^3 and re, im of N*N*(N/2+1) values");
    return NULL;
  }
  int rc = nsl_spectral_forward((int32_t)N, f, re, im);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

static napi_value js_spectral_inverse(napi_env env, napi_callback_info info) {
  ARGS(4);
  int64_t N = 0;
  double *f, *re, *im;
  size_t n1, n2, n3;
  get_i64(env, argv[0], &N);
  if (!get_f64arr(env, argv[1], &re, &n1) || !get_f64arr(env, argv[2], &im, &n2) || !get_f64arr(env, argv[3], &f, &n3)) {
    napi_throw_type_error(env, NULL, "spectralInverse(N, re, im, f)");
    return NULL;
  }
  if (!spectral_sizes_ok(N, n3, n1, n2)) {
    napi_throw_range_error(env, NULL, "spectralInverse needs re, im of N*N*(N/2+1) and f of N^3 values");
    return NULL;
  }
  int rc = nsl_spectral_inverse((int32_t)N, re, im, f);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

static napi_value js_factorize(napi_env env, napi_callback_info info) {
  ARGS(1);
  int64_t N = 0;
  get_i64(env, argv[0], &N);
  int32_t f[64];
  int k = nsl_factorize((int32_t)N, f, 64);
  if (k < 0) return throw_status(env, k);
  napi_value arr;
  CALL(env, napi_create_array_with_length(env, (size_t)k, &arr));
  for (int i = 0; i < k; i++) {
    napi_value v;
    CALL(env, napi_create_int32(env, f[i], &v));
    CALL(env, napi_set_element(env, arr, (uint32_t)i, v));
  }
  return arr;
}

static napi_value js_profile(napi_env env, napi_callback_info info) {
  ARGS(2);
  SOLVER(0);
  int64_t on = 0;
  get_i64(env, argv[1], &on);
  int rc = nsl_profile(s, (int32_t)on);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

static napi_value js_profile_read(napi_env env, napi_callback_info info) {
  ARGS(1);
  SOLVER(0);
  double ms[NSL_PROF_CLASSES];
  int64_t cnt[NSL_PROF_CLASSES];
  int rc = nsl_profile_read(s, ms, cnt);
  if (rc < 0) return throw_status(env, rc);
  double *p;
  napi_value out = new_f64arr(env, 2 * NSL_PROF_CLASSES, &p);
  for (int i = 0; i < NSL_PROF_CLASSES; i++) { p[i] = ms[i]; p[NSL_PROF_CLASSES + i] = (double)cnt[i]; }
  return out;
}

/* ------------------------------------------------------------------------------------------------ pools */
/* A pool handle: freed by poolDestroy or, failing that, when the handle is collected. */
typedef struct {
  nsl_pool *p;
} pool_ref;

static void finalize_pool(napi_env env, void *data, void *hint) {
  (void)env; (void)hint;
  pool_ref *r = (pool_ref *)data;
  if (r->p) nsl_pool_destroy(r->p);
  free(r);
}

static nsl_pool *get_pool(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || p == NULL) return NULL;
  return ((pool_ref *)p)->p;
}

#define POOL(idx)                                                           \
  nsl_pool *pl = get_pool(env, argv[idx]);                                  \
  if (!pl) { napi_throw_type_error(env, NULL, "not a pool handle"); return NULL; }

static napi_value js_pool_slot_bytes(napi_env env, napi_callback_info info) {
  ARGS(1);
  int64_t N = 0;
  get_i64(env, argv[0], &N);
  const int64_t b = nsl_pool_slot_bytes((int32_t)N);
  if (b < 0) return throw_status(env, (int)b);
  napi_value v;
  CALL(env, napi_create_double(env, (double)b, &v));
  return v;
}

/* poolCreate(N, slots) -> pool handle (destroyed when collected) */
static napi_value js_pool_create(napi_env env, napi_callback_info info) {
  ARGS(2);
  int64_t N = 0, slots = 0;
  get_i64(env, argv[0], &N);
  get_i64(env, argv[1], &slots);
  nsl_pool *p = NULL;
  int rc = nsl_pool_create((int32_t)N, (int32_t)slots, &p);
  if (rc < 0) return throw_status(env, rc);
  pool_ref *r = (pool_ref *)malloc(sizeof(pool_ref));
  if (!r) {
    nsl_pool_destroy(p);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  r->p = p;
  napi_value ext;
  CALL(env, napi_create_external(env, r, finalize_pool, NULL, &ext));
  return ext;
}

/* poolDestroy(p): frees the pool's device memory now. The pool's slot handles must not be used afterwards; the pool
 * handle then throws "not a pool handle". */
static napi_value js_pool_destroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  void *d = NULL;
  if (napi_get_value_external(env, argv[0], &d) != napi_ok || d == NULL) {
    napi_throw_type_error(env, NULL, "not a pool handle");
    return NULL;
  }
  pool_ref *r = (pool_ref *)d;
  if (r->p) nsl_pool_destroy(r->p);
  r->p = NULL;
  return NULL;
}

/* A group of pools (nsl_group_*). The JS side keeps the pools referenced for the group's life and destroys the
 * group before any of its pools. */
typedef struct {
  nsl_group *g;
} group_ref;

static void finalize_group(napi_env env, void *data, void *hint) {
  (void)env; (void)hint;
  group_ref *r = (group_ref *)data;
  if (r->g) nsl_group_destroy(r->g);
  free(r);
}

/* groupCreate([pool handles]) -> group handle */
static napi_value js_group_create(napi_env env, napi_callback_info info) {
  ARGS(1);
  uint32_t n = 0;
  if (napi_get_array_length(env, argv[0], &n) != napi_ok || n == 0 || n > 1024) {
    napi_throw_type_error(env, NULL, "groupCreate([pool, ...])");
    return NULL;
  }
  nsl_pool *pools[1024];
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    CALL(env, napi_get_element(env, argv[0], i, &e));
    pools[i] = get_pool(env, e);
    if (!pools[i]) { napi_throw_type_error(env, NULL, "not a pool handle"); return NULL; }
  }
  nsl_group *g = NULL;
  int rc = nsl_group_create(pools, (int32_t)n, &g);
  if (rc < 0) return throw_status(env, rc);
  group_ref *r = (group_ref *)malloc(sizeof(group_ref));
  if (!r) { nsl_group_destroy(g); napi_throw_error(env, NULL, "out of host memory"); return NULL; }
  r->g = g;
  napi_value ext;
  CALL(env, napi_create_external(env, r, finalize_group, NULL, &ext));
  return ext;
}

static group_ref *get_group(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || p == NULL) return NULL;
  return (group_ref *)p;
}

/* groupLaunch(g, k) */
static napi_value js_group_launch(napi_env env, napi_callback_info info) {
  ARGS(2);
  group_ref *r = get_group(env, argv[0]);
  if (!r || !r->g) { napi_throw_type_error(env, NULL, "not a group handle"); return NULL; }
  int64_t k = 64;
  get_i64(env, argv[1], &k);
  int rc = nsl_group_launch(r->g, (int32_t)k);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

/* groupDestroy(g): frees the group's graph now (its pools stay) */
static napi_value js_group_destroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  group_ref *r = get_group(env, argv[0]);
  if (!r) { napi_throw_type_error(env, NULL, "not a group handle"); return NULL; }
  if (r->g) nsl_group_destroy(r->g);
  r->g = NULL;
  return NULL;
}

/* poolInfo(p) -> [N, slots, used, graphed, bytes] */
static napi_value js_pool_info(napi_env env, napi_callback_info info) {
  ARGS(1);
  POOL(0);
  int32_t N, slots, used, graphed;
  int64_t bytes;
  int rc = nsl_pool_info(pl, &N, &slots, &used, &graphed, &bytes);
  if (rc < 0) return throw_status(env, rc);
  double *d;
  napi_value out = new_f64arr(env, 5, &d);
  d[0] = N; d[1] = slots; d[2] = used; d[3] = graphed; d[4] = (double)bytes;
  return out;
}

/* poolAdmit(p, N, ic, nu, cfl, dt, icp0..icp3) -> [slot, solver handle]. The handle is the slot's view: it has no
 * finaliser (nsl_destroy ignores views); the slot is freed by poolRelease and the memory with the pool. */
static napi_value js_pool_admit(napi_env env, napi_callback_info info) {
  ARGS(10);
  POOL(0);
  nsl_opts o;
  memset(&o, 0, sizeof o);
  int64_t N = 0, ic = 0;
  if (argc < 10 || !get_i64(env, argv[1], &N) || !get_i64(env, argv[2], &ic) || !get_f64(env, argv[3], &o.nu) ||
      !get_f64(env, argv[4], &o.cfl) || !get_f64(env, argv[5], &o.dt)) {
    napi_throw_type_error(env, NULL, "poolAdmit(p, N, ic, nu, cfl, dt, p0, p1, p2, p3)");
    return NULL;
  }
  for (int i = 0; i < 4; i++) if (!get_f64(env, argv[6 + i], &o.icp[i])) o.icp[i] = 0;
  o.N = (int32_t)N;
  o.ic = (int32_t)ic;
  int32_t slot = -1;
  int rc = nsl_pool_admit(pl, &o, &slot);
  if (rc < 0) return throw_status(env, rc);
  nsl_solver *v = NULL;
  rc = nsl_pool_slot(pl, slot, &v);
  if (rc < 0) return throw_status(env, rc);
  napi_value arr, sv, ext = wrap_solver(env, v, 0);
  if (ext == NULL) return NULL;
  CALL(env, napi_create_array_with_length(env, 2, &arr));
  CALL(env, napi_create_int32(env, slot, &sv));
  CALL(env, napi_set_element(env, arr, 0, sv));
  CALL(env, napi_set_element(env, arr, 1, ext));
  return arr;
}

static napi_value js_pool_set_target(napi_env env, napi_callback_info info) {
  ARGS(5);
  POOL(0);
  int64_t slot = 0, n = 0, has = 0;
  double tend = 0;
  get_i64(env, argv[1], &slot);
  get_i64(env, argv[2], &n);
  get_i64(env, argv[3], &has);
  get_f64(env, argv[4], &tend);
  int rc = nsl_pool_set_target(pl, (int32_t)slot, n, (int32_t)has, tend);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

static napi_value js_pool_launch(napi_env env, napi_callback_info info) {
  ARGS(2);
  POOL(0);
  int64_t k = 0;
  get_i64(env, argv[1], &k);
  int rc = nsl_pool_launch(pl, (int32_t)k);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

/* poolCollect(p, k) -> Float64Array: [running, done[0..slots), then slots*k records of 21 values
 * (t dt step ebal zbal last[8] d0[8]); slot i's records start at 1 + slots + i*k*21] */
static napi_value js_pool_collect(napi_env env, napi_callback_info info) {
  ARGS(2);
  POOL(0);
  int64_t k = 0;
  get_i64(env, argv[1], &k);
  int32_t N, slots, used, graphed;
  int64_t bytes;
  int rc = nsl_pool_info(pl, &N, &slots, &used, &graphed, &bytes);
  if (rc < 0) return throw_status(env, rc);
  if (k < 1) k = 1;
  nsl_steprec *recs = (nsl_steprec *)calloc((size_t)slots * (size_t)k, sizeof(nsl_steprec));
  int64_t *done = (int64_t *)calloc((size_t)slots, sizeof(int64_t));
  if (!recs || !done) { free(recs); free(done); napi_throw_error(env, NULL, "out of host memory"); return NULL; }
  int32_t running = 0;
  rc = nsl_pool_collect(pl, recs, (int32_t)k, done, &running);
  if (rc < 0) { free(recs); free(done); return throw_status(env, rc); }
  double *d;
  const size_t head = 1 + (size_t)slots;
  napi_value out = new_f64arr(env, head + (size_t)slots * (size_t)k * 21, &d);
  d[0] = running;
  for (int32_t i = 0; i < slots; i++) {
    d[1 + i] = (double)done[i];
    for (int64_t j = 0; j < done[i] && j < k; j++) {
      const nsl_steprec *r = &recs[(size_t)i * (size_t)k + (size_t)j];
      double *o = d + head + ((size_t)i * (size_t)k + (size_t)j) * 21;
      o[0] = r->t; o[1] = r->dt; o[2] = (double)r->step; o[3] = r->ebal; o[4] = r->zbal;
      memcpy(o + 5, &r->last, 8 * sizeof(double));
      memcpy(o + 13, &r->d0, 8 * sizeof(double));
    }
  }
  free(recs);
  free(done);
  return out;
}

static napi_value js_pool_ready(napi_env env, napi_callback_info info) {
  ARGS(1);
  POOL(0);
  int rc = nsl_pool_ready(pl);
  if (rc < 0) return throw_status(env, rc);
  napi_value v;
  CALL(env, napi_get_boolean(env, rc == 1, &v));
  return v;
}

/* deviceMemory() -> [free, total] bytes */
static napi_value js_device_memory(napi_env env, napi_callback_info info) {
  (void)info;
  int64_t f = 0, t = 0;
  int rc = nsl_device_memory(&f, &t);
  if (rc < 0) return throw_status(env, rc);
  double *d;
  napi_value out = new_f64arr(env, 2, &d);
  d[0] = (double)f; d[1] = (double)t;
  return out;
}

static napi_value js_pool_release(napi_env env, napi_callback_info info) {
  ARGS(2);
  POOL(0);
  int64_t slot = 0;
  get_i64(env, argv[1], &slot);
  int rc = nsl_pool_release(pl, (int32_t)slot);
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

/* tune forms as [xpass, lp, xwaves, zfused] */
static napi_value tune_arr(napi_env env, const nsl_tune *t) {
  napi_value arr;
  const int32_t v7[7] = {t->xpass, t->lp, t->xwaves, t->zfused, t->ptile, t->xsched, t->zsched};
  CALL(env, napi_create_array_with_length(env, 7, &arr));
  for (uint32_t i = 0; i < 7; i++) {
    napi_value v;
    CALL(env, napi_create_int32(env, v7[i], &v));
    CALL(env, napi_set_element(env, arr, i, v));
  }
  return arr;
}

/* tuneGet(N) -> [xpass, lp, xwaves, zfused, ptile, xsched, zsched] */
static napi_value js_tune_get(napi_env env, napi_callback_info info) {
  ARGS(1);
  int64_t N = 0;
  get_i64(env, argv[0], &N);
  nsl_tune t;
  int rc = nsl_tune_get((int32_t)N, &t);
  if (rc < 0) return throw_status(env, rc);
  return tune_arr(env, &t);
}

/* tuneSet(N, xpass, lp, xwaves, zfused[, ptile, xsched, zsched]), or tuneSet(N) to drop the override */
static napi_value js_tune_set(napi_env env, napi_callback_info info) {
  ARGS(8);
  int64_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < argc && i < 8; i++)
    if (!get_i64(env, argv[i], &a[i])) a[i] = 0;
  int rc;
  if (argc < 5) {
    rc = nsl_tune_set((int32_t)a[0], NULL);
  } else {
    nsl_tune t = {(int32_t)a[1], (int32_t)a[2], (int32_t)a[3], (int32_t)a[4], (int32_t)a[5], (int32_t)a[6], (int32_t)a[7]};
    rc = nsl_tune_set((int32_t)a[0], &t);
  }
  if (rc < 0) return throw_status(env, rc);
  return NULL;
}

/* tuneCandidates(N) -> [[xpass, lp, xwaves, zfused, ptile, xsched, zsched], ...] */
static napi_value js_tune_candidates(napi_env env, napi_callback_info info) {
  ARGS(1);
  int64_t N = 0;
  get_i64(env, argv[0], &N);
  nsl_tune t[512];
  int k = nsl_tune_candidates((int32_t)N, t, 512);
  if (k < 0) return throw_status(env, k);
  if (k > 512) k = 512;
  napi_value arr;
  CALL(env, napi_create_array_with_length(env, (size_t)k, &arr));
  for (int i = 0; i < k; i++) {
    napi_value e = tune_arr(env, &t[i]);
    if (e == NULL) return NULL;
    CALL(env, napi_set_element(env, arr, (uint32_t)i, e));
  }
  return arr;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  CALL(env, napi_create_string_utf8(env, nsl_version(), NAPI_AUTO_LENGTH, &v));
  return v;
}

static napi_value Init(napi_env env, napi_value exports) {
  static const struct { const char *name; napi_callback cb; } fns[] = {
    {"create", js_create}, {"destroy", js_destroy}, {"setIC", js_set_ic}, {"run", js_run}, {"last", js_last},
    {"setClock", js_set_clock}, {"getClock", js_get_clock}, {"norms", js_norms},
    {"getState", js_get_state}, {"setState", js_set_state}, {"diagnose", js_diagnose},
    {"spectrum", js_spectrum}, {"interpMax", js_interp_max}, {"field", js_field},
    {"exactError", js_exact_error}, {"deviceBytes", js_device_bytes}, {"fft1", js_fft1},
    {"spectralForward", js_spectral_forward}, {"spectralInverse", js_spectral_inverse},
    {"factorize", js_factorize}, {"profile", js_profile}, {"profileRead", js_profile_read},
    {"version", js_version}, {"poolSlotBytes", js_pool_slot_bytes}, {"poolCreate", js_pool_create},
    {"poolInfo", js_pool_info}, {"poolAdmit", js_pool_admit}, {"poolSetTarget", js_pool_set_target},
    {"poolLaunch", js_pool_launch}, {"poolCollect", js_pool_collect}, {"poolRelease", js_pool_release},
    {"poolDestroy", js_pool_destroy},
    {"deviceMemory", js_device_memory}, {"poolReady", js_pool_ready},
    {"tuneGet", js_tune_get}, {"tuneSet", js_tune_set}, {"tuneCandidates", js_tune_candidates},
    {"slice", js_slice}, {"imageDiag", js_image_diag}, {"cutoffPileUp", js_cutoff_pile_up},
    {"groupCreate", js_group_create}, {"groupLaunch", js_group_launch}, {"groupDestroy", js_group_destroy},
    {"precision", js_precision},
  };
  for (size_t i = 0; i < sizeof fns / sizeof fns[0]; i++) {
    napi_value f;
    CALL(env, napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].cb, NULL, &f));
    CALL(env, napi_set_named_property(env, exports, fns[i].name, f));
  }
  return exports;
}

NAPI_MODULE(nslab_napi, Init)
