/*
 * ORACLE — test infrastructure only. A single-threaded CPU restatement of the NSLab pseudo-spectral Navier–Stokes
 * core, used by tests/ (as the parity checker), by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 * Nothing in the product library links, loads or calls this file.
 *
 * It follows the reference algorithm step by step, including the loop orders that fix summation order and the
 * PRNG draw sequence:
 *   factorize / fftPlan / fft1 ............ pocket-wind-tunnel/src/nslab.js:28-92
 *   Spectral.forward / inverse ............ nslab.js:94-155
 *   solver set-up (kc, kx, kw, mask, hw) .. nslab.js:171-191
 *   project / curlHat / norms / rhs ....... nslab.js:194-247
 *   setIC (tgv, tgv2d, abc, tubes, random)  nslab.js:161, 250-302
 *   tubesz (no reference counterpart) ..... the memory-heavy benchmark's tube pair along z: the tubes path of
 *                                           nslab.js:257-272 on another vorticity field, plus the random field of
 *                                           nslab.js:273-286 scaled to the perturbation energy
 *   chooseDt / step / run ................. nslab.js:305-334
 *   diagnose / eig3 / spectrum ............ nslab.js:339-420
 *   interpMax / imageDiag / cutoffPileUp .. nslab.js:421-564
 *   slice / exactError .................... nslab.js:566-594
 * Pinned by tests/test_oracle.py against the reference's golden vectors (tests/golden/) and closed forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TWO_PI (2.0 * M_PI)

/* ---------------------------------------------------------------- FFT (nslab.js:28-92) */
typedef struct { int r, n, m; double *wc, *ws; } or_stage;
typedef struct { int N, nst; or_stage st[40]; double *xr, *xi, *yr, *yi; } or_plan;

int or_factorize(int N, int *f) {
  int n = N, k = 0;
  while (n % 4 == 0) { f[k++] = 4; n /= 4; }
  while (n % 2 == 0) { f[k++] = 2; n /= 2; }
  while (n % 3 == 0) { f[k++] = 3; n /= 3; }
  if (n != 1 || N < 4) return -1;
  return k;
}

or_plan *or_plan_create(int N) {
  int f[40], k = or_factorize(N, f);
  if (k < 0) return NULL;
  or_plan *p = (or_plan *)calloc(1, sizeof(or_plan));
  p->N = N; p->nst = k;
  int n = N;
  for (int s = 0; s < k; s++) {
    int r = f[s], m = n / r;
    or_stage *S = &p->st[s];
    S->r = r; S->n = n; S->m = m;
    S->wc = (double *)malloc(sizeof(double) * (r - 1) * m);
    S->ws = (double *)malloc(sizeof(double) * (r - 1) * m);
    for (int j = 1; j < r; j++)
      for (int q = 0; q < m; q++) {
        double a = TWO_PI * j * q / n;
        S->wc[(j - 1) * m + q] = cos(a); S->ws[(j - 1) * m + q] = sin(a);
      }
    n = m;
  }
  p->xr = (double *)calloc(N, 8); p->xi = (double *)calloc(N, 8);
  p->yr = (double *)calloc(N, 8); p->yi = (double *)calloc(N, 8);
  return p;
}

void or_plan_free(or_plan *p) {
  if (!p) return;
  for (int s = 0; s < p->nst; s++) { free(p->st[s].wc); free(p->st[s].ws); }
  free(p->xr); free(p->xi); free(p->yr); free(p->yi); free(p);
}

/* Transforms p->xr/xi; the result is left in outr, outi (one of the two ping-pong buffers). */
static void fft1(or_plan *pl, int sign, double **outr, double **outi) {
  double *xr = pl->xr, *xi = pl->xi, *yr = pl->yr, *yi = pl->yi, *t;
  int s = 1;
  const double S3 = -sign * 0.8660254037844386;
  for (int g = 0; g < pl->nst; g++) {
    const or_stage *st = &pl->st[g];
    const int r = st->r, m = st->m;
    const double *wc = st->wc, *ws = st->ws;
    if (r == 2) {
      for (int p = 0; p < m; p++) {
        const double c = wc[p], d = sign * ws[p];
        for (int q = 0; q < s; q++) {
          int i0 = q + s * p, i1 = i0 + s * m, o0 = q + s * 2 * p, o1 = o0 + s;
          double ar = xr[i0], ai = xi[i0], br = xr[i1], bi = xi[i1];
          yr[o0] = ar + br; yi[o0] = ai + bi;
          double tr = ar - br, ti = ai - bi;
          yr[o1] = tr * c - ti * d; yi[o1] = tr * d + ti * c;
        }
      }
    } else if (r == 4) {
      for (int p = 0; p < m; p++) {
        const double c1 = wc[p], d1 = sign * ws[p], c2 = wc[m + p], d2 = sign * ws[m + p], c3 = wc[2 * m + p], d3 = sign * ws[2 * m + p];
        for (int q = 0; q < s; q++) {
          int i0 = q + s * p, i1 = i0 + s * m, i2 = i1 + s * m, i3 = i2 + s * m;
          int o0 = q + s * 4 * p, o1 = o0 + s, o2 = o1 + s, o3 = o2 + s;
          double a0r = xr[i0], a0i = xi[i0], a1r = xr[i1], a1i = xi[i1], a2r = xr[i2], a2i = xi[i2], a3r = xr[i3], a3i = xi[i3];
          double t0r = a0r + a2r, t0i = a0i + a2i, t1r = a0r - a2r, t1i = a0i - a2i;
          double t2r = a1r + a3r, t2i = a1i + a3i, t3r = a1r - a3r, t3i = a1i - a3i;
          yr[o0] = t0r + t2r; yi[o0] = t0i + t2i;
          double br = t1r - sign * t3i, bi = t1i + sign * t3r;
          yr[o1] = br * c1 - bi * d1; yi[o1] = br * d1 + bi * c1;
          br = t0r - t2r; bi = t0i - t2i;
          yr[o2] = br * c2 - bi * d2; yi[o2] = br * d2 + bi * c2;
          br = t1r + sign * t3i; bi = t1i - sign * t3r;
          yr[o3] = br * c3 - bi * d3; yi[o3] = br * d3 + bi * c3;
        }
      }
    } else {
      for (int p = 0; p < m; p++) {
        const double c1 = wc[p], d1 = sign * ws[p], c2 = wc[m + p], d2 = sign * ws[m + p];
        for (int q = 0; q < s; q++) {
          int i0 = q + s * p, i1 = i0 + s * m, i2 = i1 + s * m, o0 = q + s * 3 * p, o1 = o0 + s, o2 = o1 + s;
          double a0r = xr[i0], a0i = xi[i0], a1r = xr[i1], a1i = xi[i1], a2r = xr[i2], a2i = xi[i2];
          double sr = a1r + a2r, si = a1i + a2i, ur = a0r - 0.5 * sr, ui = a0i - 0.5 * si;
          double vr = S3 * (a1r - a2r), vi = S3 * (a1i - a2i);
          yr[o0] = a0r + sr; yi[o0] = a0i + si;
          double br = ur + vi, bi = ui - vr;
          yr[o1] = br * c1 - bi * d1; yi[o1] = br * d1 + bi * c1;
          br = ur - vi; bi = ui + vr;
          yr[o2] = br * c2 - bi * d2; yi[o2] = br * d2 + bi * c2;
        }
      }
    }
    t = xr; xr = yr; yr = t; t = xi; xi = yi; yi = t; s *= r;
  }
  *outr = xr; *outi = xi;
}

/* One line, in place on (re, im) of length N: the exported fft1 (plan.xr/xi in, [re, im] out). */
int or_fft1(int N, int sign, double *re, double *im) {
  or_plan *p = or_plan_create(N);
  if (!p) return -1;
  memcpy(p->xr, re, 8 * N); memcpy(p->xi, im, 8 * N);
  double *Zr, *Zi; fft1(p, sign, &Zr, &Zi);
  memcpy(re, Zr, 8 * N); memcpy(im, Zi, 8 * N);
  or_plan_free(p);
  return 0;
}

/* ---------------------------------------------------------------- 3D transforms (nslab.js:94-155) */
typedef struct { int N, NH; long NS, N3; or_plan *plan; double *wr, *wi; } or_spectral;

static or_spectral *spectral_create(int N) {
  or_spectral *sp = (or_spectral *)calloc(1, sizeof(or_spectral));
  sp->N = N; sp->NH = N / 2 + 1; sp->NS = (long)N * N * sp->NH; sp->N3 = (long)N * N * N;
  sp->plan = or_plan_create(N);
  sp->wr = (double *)calloc(sp->NS, 8); sp->wi = (double *)calloc(sp->NS, 8);
  return sp;
}
static void spectral_free(or_spectral *sp) { or_plan_free(sp->plan); free(sp->wr); free(sp->wi); free(sp); }

static void spectral_forward(or_spectral *sp, const double *f, double *re, double *im) {
  const int N = sp->N, NH = sp->NH, half = N / 2;
  double *xr = sp->plan->xr, *xi = sp->plan->xi, *Zr, *Zi;
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y += 2) {
      long b0 = ((long)z * N + y) * N, b1 = b0 + N;
      for (int x = 0; x < N; x++) { xr[x] = f[b0 + x]; xi[x] = f[b1 + x]; }
      fft1(sp->plan, -1, &Zr, &Zi);
      long o0 = ((long)z * N + y) * NH, o1 = o0 + NH;
      for (int k = 0; k <= half; k++) {
        int kk = (N - k) % N;
        double zr = Zr[k], zi = Zi[k], cr = Zr[kk], ci = -Zi[kk];
        re[o0 + k] = 0.5 * (zr + cr); im[o0 + k] = 0.5 * (zi + ci);
        re[o1 + k] = 0.5 * (zi - ci); im[o1 + k] = -0.5 * (zr - cr);
      }
    }
  for (int z = 0; z < N; z++)
    for (int k = 0; k < NH; k++) {
      long base = (long)z * N * NH + k;
      for (int j = 0; j < N; j++) { long idx = base + (long)j * NH; xr[j] = re[idx]; xi[j] = im[idx]; }
      fft1(sp->plan, -1, &Zr, &Zi);
      for (int j = 0; j < N; j++) { long idx = base + (long)j * NH; re[idx] = Zr[j]; im[idx] = Zi[j]; }
    }
  const double inv = 1.0 / (double)sp->N3;
  const long stride = (long)N * NH;
  for (int j = 0; j < N; j++)
    for (int k = 0; k < NH; k++) {
      long base = (long)j * NH + k;
      for (int l = 0; l < N; l++) { long idx = base + l * stride; xr[l] = re[idx]; xi[l] = im[idx]; }
      fft1(sp->plan, -1, &Zr, &Zi);
      for (int l = 0; l < N; l++) { long idx = base + l * stride; re[idx] = Zr[l] * inv; im[idx] = Zi[l] * inv; }
    }
}

static void spectral_inverse(or_spectral *sp, const double *re, const double *im, double *f) {
  const int N = sp->N, NH = sp->NH, half = N / 2;
  const long stride = (long)N * NH;
  double *xr = sp->plan->xr, *xi = sp->plan->xi, *wr = sp->wr, *wi = sp->wi, *Zr, *Zi;
  for (int j = 0; j < N; j++)
    for (int k = 0; k < NH; k++) {
      long base = (long)j * NH + k;
      for (int l = 0; l < N; l++) { long idx = base + l * stride; xr[l] = re[idx]; xi[l] = im[idx]; }
      fft1(sp->plan, 1, &Zr, &Zi);
      for (int l = 0; l < N; l++) { long idx = base + l * stride; wr[idx] = Zr[l]; wi[idx] = Zi[l]; }
    }
  for (int z = 0; z < N; z++)
    for (int k = 0; k < NH; k++) {
      long base = (long)z * N * NH + k;
      for (int j = 0; j < N; j++) { long idx = base + (long)j * NH; xr[j] = wr[idx]; xi[j] = wi[idx]; }
      fft1(sp->plan, 1, &Zr, &Zi);
      for (int j = 0; j < N; j++) { long idx = base + (long)j * NH; wr[idx] = Zr[j]; wi[idx] = Zi[j]; }
    }
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y += 2) {
      long o0 = ((long)z * N + y) * NH, o1 = o0 + NH;
      for (int k = 0; k <= half; k++) {
        double Ar = wr[o0 + k], Ai = wi[o0 + k], Br = wr[o1 + k], Bi = wi[o1 + k];
        xr[k] = Ar - Bi; xi[k] = Ai + Br;
        if (k > 0 && k < half) { xr[N - k] = Ar + Bi; xi[N - k] = Br - Ai; }
      }
      fft1(sp->plan, 1, &Zr, &Zi);
      long b0 = ((long)z * N + y) * N, b1 = b0 + N;
      for (int x = 0; x < N; x++) { f[b0 + x] = Zr[x]; f[b1 + x] = Zi[x]; }
    }
}

int or_spectral_forward(int N, const double *f, double *re, double *im) {
  if (N % 2) return -1;
  or_spectral *sp = spectral_create(N);
  if (!sp->plan) { spectral_free(sp); return -1; }
  spectral_forward(sp, f, re, im); spectral_free(sp); return 0;
}
int or_spectral_inverse(int N, const double *re, const double *im, double *f) {
  if (N % 2) return -1;
  or_spectral *sp = spectral_create(N);
  if (!sp->plan) { spectral_free(sp); return -1; }
  spectral_inverse(sp, re, im, f); spectral_free(sp); return 0;
}

/* ---------------------------------------------------------------- solver (nslab.js:171-334) */
enum { IC_TGV = 0, IC_TGV2D = 1, IC_ABC = 2, IC_TUBES = 3, IC_RANDOM = 4, IC_TUBESZ = 5 };

typedef struct { double E, Z, P, eps, Tnl, Pspec, omMax, uMax; } or_diag;

typedef struct {
  double t, dt, E0, Z0, maxEbal, maxZbal, maxTnl, divMax;
  long step;
} or_state;

typedef struct {
  int N, NH, kc; long NS, N3;
  double nu, cfl, dtFixed;
  or_spectral *sp;
  double *kx, *kw, *hw; uint8_t *mask;
  double *S[6], *T[6], *K[6], *ACC[6], *NL[6], *up[3], *om[3];
  double *G[3], *Sxx, *Syy, *Szz, *Sxy, *Sxz, *Syz, *G2, *PST, *IW[6];
  or_diag last;
  or_state st;
  int ic; double icp[4];
  /* series ring (t, E, Z, eps, omMax, uMax, dt, ebal, zbal, Pspec, Pal) kept by the caller */
  double ebal, zbal;
} or_solver;

static double *alloc0(long n) { return (double *)calloc(n, 8); }

static void project(or_solver *s, double **U, int applyMask) {
  const int N = s->N, NH = s->NH;
  double *ar = U[0], *ai = U[1], *br = U[2], *bi = U[3], *cr = U[4], *ci = U[5];
  for (int l = 0; l < N; l++) { const double kz = s->kw[l];
    for (int j = 0; j < N; j++) { const double ky = s->kw[j]; const long row = ((long)l * N + j) * NH;
      for (int i = 0; i < NH; i++) {
        const long idx = row + i; const double kxx = s->kx[i], k2 = kxx * kxx + ky * ky + kz * kz;
        if (k2 == 0 || (applyMask && !s->mask[idx])) { ar[idx] = ai[idx] = br[idx] = bi[idx] = cr[idx] = ci[idx] = 0; continue; }
        const double dr = (kxx * ar[idx] + ky * br[idx] + kz * cr[idx]) / k2, di = (kxx * ai[idx] + ky * bi[idx] + kz * ci[idx]) / k2;
        ar[idx] -= kxx * dr; ai[idx] -= kxx * di; br[idx] -= ky * dr; bi[idx] -= ky * di; cr[idx] -= kz * dr; ci[idx] -= kz * di;
      } } }
}

static void curlHat(or_solver *s, double **U, double **W) {
  const int N = s->N, NH = s->NH;
  double *ar = U[0], *ai = U[1], *br = U[2], *bi = U[3], *cr = U[4], *ci = U[5];
  for (int l = 0; l < N; l++) { const double kz = s->kw[l];
    for (int j = 0; j < N; j++) { const double ky = s->kw[j]; const long row = ((long)l * N + j) * NH;
      for (int i = 0; i < NH; i++) {
        const long idx = row + i; const double kxx = s->kx[i];
        double pr = ky * cr[idx] - kz * br[idx], pi = ky * ci[idx] - kz * bi[idx]; W[0][idx] = -pi; W[1][idx] = pr;
        pr = kz * ar[idx] - kxx * cr[idx]; pi = kz * ai[idx] - kxx * ci[idx]; W[2][idx] = -pi; W[3][idx] = pr;
        pr = kxx * br[idx] - ky * ar[idx]; pi = kxx * bi[idx] - ky * ai[idx]; W[4][idx] = -pi; W[5][idx] = pr;
      } } }
}

static void norms(or_solver *s, double **U, double *E_, double *Z_, double *P_) {
  const int N = s->N, NH = s->NH;
  double E = 0, Z = 0, P = 0;
  for (int l = 0; l < N; l++) { const double kz = s->kw[l];
    for (int j = 0; j < N; j++) { const double ky = s->kw[j]; const long row = ((long)l * N + j) * NH;
      for (int i = 0; i < NH; i++) {
        const long idx = row + i; const double k2 = s->kx[i] * s->kx[i] + ky * ky + kz * kz;
        const double e = s->hw[i] * (U[0][idx] * U[0][idx] + U[1][idx] * U[1][idx] + U[2][idx] * U[2][idx] + U[3][idx] * U[3][idx] + U[4][idx] * U[4][idx] + U[5][idx] * U[5][idx]);
        E += e; Z += k2 * e; P += k2 * k2 * e;
      } } }
  *E_ = 0.5 * E; *Z_ = 0.5 * Z; *P_ = 0.5 * P;
}

static void rhs(or_solver *s, double **U, double **OUT) {
  const int N = s->N, NH = s->NH; const long N3 = s->N3; const double nu = s->nu;
  curlHat(s, U, s->NL);
  for (int c = 0; c < 3; c++) { spectral_inverse(s->sp, U[2 * c], U[2 * c + 1], s->up[c]); spectral_inverse(s->sp, s->NL[2 * c], s->NL[2 * c + 1], s->om[c]); }
  double omMax = 0, uMax = 0;
  double *u = s->up[0], *v = s->up[1], *w = s->up[2], *ox = s->om[0], *oy = s->om[1], *oz = s->om[2];
  for (long n = 0; n < N3; n++) {
    const double a = u[n], b = v[n], c = w[n], p = ox[n], q = oy[n], r = oz[n];
    const double o2 = p * p + q * q + r * r, u2 = a * a + b * b + c * c;
    if (o2 > omMax) omMax = o2;
    if (u2 > uMax) uMax = u2;
    ox[n] = b * r - c * q; oy[n] = c * p - a * r; oz[n] = a * q - b * p;
  }
  for (int c = 0; c < 3; c++) spectral_forward(s->sp, s->om[c], s->NL[2 * c], s->NL[2 * c + 1]);
  project(s, s->NL, 1);
  double Tnl = 0, Pspec = 0, E = 0, Z = 0, Pal = 0;
  for (int l = 0; l < N; l++) { const double kz = s->kw[l];
    for (int j = 0; j < N; j++) { const double ky = s->kw[j]; const long row = ((long)l * N + j) * NH;
      for (int i = 0; i < NH; i++) {
        const long idx = row + i; const double k2 = s->kx[i] * s->kx[i] + ky * ky + kz * kz, h = s->hw[i];
        double dot = 0, e = 0;
        for (int c = 0; c < 3; c++) {
          const double ur = U[2 * c][idx], ui = U[2 * c + 1][idx], nr = s->NL[2 * c][idx], ni = s->NL[2 * c + 1][idx];
          dot += ur * nr + ui * ni; e += ur * ur + ui * ui;
          OUT[2 * c][idx] = nr - nu * k2 * ur; OUT[2 * c + 1][idx] = ni - nu * k2 * ui;
        }
        Tnl += h * dot; Pspec += h * k2 * dot; E += h * e; Z += h * k2 * e; Pal += h * k2 * k2 * e;
      } } }
  s->last.E = 0.5 * E; s->last.Z = 0.5 * Z; s->last.P = 0.5 * Pal; s->last.eps = nu * Z;
  s->last.Tnl = Tnl; s->last.Pspec = Pspec; s->last.omMax = sqrt(omMax); s->last.uMax = sqrt(uMax);
}

/* Mulberry32 (nslab.js:161): a 32-bit counter-hash PRNG. */
typedef struct { uint32_t a; } mb32;
static double mb32_next(mb32 *g) {
  g->a = g->a + 0x6D2B79F5u;
  uint32_t a = g->a;
  uint32_t t = (a ^ (a >> 15)) * (1u | a);
  t = (t + (t ^ (t >> 7)) * (61u | t)) ^ t;
  return (double)(t ^ (t >> 14)) / 4294967296.0;
}

static void finishIC(or_solver *s) {
  s->st.maxEbal = s->st.maxZbal = s->st.maxTnl = s->st.divMax = 0;
  rhs(s, s->S, s->K);
  s->st.E0 = s->last.E; s->st.Z0 = s->last.Z;
}

/* Random solenoidal field (nslab.js:273-286) into dst: Mulberry32(seed) Box-Muller draws over the masked modes,
 * Hermitian fill of kx = 0, project, scale to energy E0. */
static void random_field(or_solver *s, double **dst, double seed, double k0, double E0) {
  const int N = s->N, NH = s->NH;
  mb32 g = { (uint32_t)(int32_t)seed };
  for (int c = 0; c < 6; c++) memset(dst[c], 0, 8 * s->NS);
  for (int l = 0; l < N; l++) { const double kz = s->kw[l];
    for (int j = 0; j < N; j++) { const double ky = s->kw[j]; const long row = ((long)l * N + j) * NH;
      for (int i = 0; i < NH; i++) {
        const long idx = row + i; if (!s->mask[idx]) continue;
        const double k2 = s->kx[i] * s->kx[i] + ky * ky + kz * kz; if (k2 == 0) continue;
        if (i == 0 && (kz < 0 || (kz == 0 && ky < 0))) continue;
        const double k = sqrt(k2), amp = sqrt(pow(k, 4) * exp(-2 * (k / k0) * (k / k0)) / (4 * M_PI * k2));
        for (int c = 0; c < 3; c++) {
          for (int part = 0; part < 2; part++) {
            double a = mb32_next(&g); a = a > 1e-12 ? a : 1e-12;
            const double b = mb32_next(&g);
            dst[2 * c + part][idx] = amp * (sqrt(-2 * log(a)) * cos(TWO_PI * b));
          }
        }
        if (i == 0) {
          const int jm = (N - j) % N, lm = (N - l) % N; const long idm = ((long)lm * N + jm) * NH;
          for (int c = 0; c < 3; c++) { dst[2 * c][idm] = dst[2 * c][idx]; dst[2 * c + 1][idm] = -dst[2 * c + 1][idx]; }
        }
      } } }
  project(s, dst, 1);
  double E, Z, P; norms(s, dst, &E, &Z, &P);
  const double sc = sqrt(E0 / E);
  for (int c = 0; c < 6; c++) for (long n = 0; n < s->NS; n++) dst[c][n] *= sc;
}

int or_setIC(or_solver *s, int kind, const double *prm) {
  const int N = s->N, NH = s->NH; const double h = TWO_PI / N;
  double *u = s->up[0], *v = s->up[1], *w = s->up[2];
  s->ic = kind; for (int q = 0; q < 4; q++) s->icp[q] = prm ? prm[q] : 0;
  if (kind == IC_TGV || kind == IC_TGV2D || kind == IC_ABC) {
    const double A = prm[0], B = prm[1], C = prm[2];
    for (int l = 0; l < N; l++) for (int j = 0; j < N; j++) for (int i = 0; i < N; i++) {
      const long n = ((long)l * N + j) * N + i; const double x = i * h, y = j * h, z = l * h;
      if (kind == IC_TGV) { u[n] = sin(x) * cos(y) * cos(z); v[n] = -cos(x) * sin(y) * cos(z); w[n] = 0; }
      else if (kind == IC_TGV2D) { u[n] = sin(x) * cos(y); v[n] = -cos(x) * sin(y); w[n] = 0; }
      else { u[n] = A * sin(z) + C * cos(y); v[n] = B * sin(x) + A * cos(z); w[n] = C * sin(y) + B * cos(x); }
    }
    for (int c = 0; c < 3; c++) spectral_forward(s->sp, s->up[c], s->S[2 * c], s->S[2 * c + 1]);
    project(s, s->S, 1);
  } else if (kind == IC_TUBES || kind == IC_TUBESZ) {
    double *ox = s->om[0], *oy = s->om[1], *oz = s->om[2];
    if (kind == IC_TUBES) {
      const double amp = prm[0], sig = prm[1], d0 = prm[2], del = prm[3];
      for (int l = 0; l < N; l++) for (int j = 0; j < N; j++) for (int i = 0; i < N; i++) {
        const long n = ((long)l * N + j) * N + i; const double x = i * h, y = j * h, z = l * h;
        const double d = d0 + del * cos(x); double sm = 0;
        for (int sgn = -1; sgn <= 1; sgn += 2) {
          const double yc = M_PI + sgn * d, zc = M_PI; double best = INFINITY;
          for (int py = -1; py <= 1; py++) for (int pz = -1; pz <= 1; pz++) {
            const double dy = y - yc + py * TWO_PI, dz = z - zc + pz * TWO_PI; const double q = dy * dy + dz * dz;
            best = q < best ? q : best;
          }
          sm += -sgn * amp * exp(-best / (sig * sig));
        }
        ox[n] = sm; oy[n] = 0; oz[n] = 0;
      }
    } else {   /* tube pair along z: cores +-Gamma/(pi sigma^2) exp(-r^2/sigma^2), centres (pi +- (d + 0.1 cos z), pi) */
      const double G = prm[0], sig = prm[1], d0 = prm[2], a = 0.1;
      for (int l = 0; l < N; l++) for (int j = 0; j < N; j++) for (int i = 0; i < N; i++) {
        const long n = ((long)l * N + j) * N + i; const double x = i * h, y = j * h, z = l * h;
        double wx = 0, wz = 0;
        for (int sgn = -1; sgn <= 1; sgn += 2) {
          const double xc = M_PI + sgn * (d0 + a * cos(z)), yc = M_PI; double best = INFINITY;
          for (int px = -1; px <= 1; px++) for (int py = -1; py <= 1; py++) {
            const double dx = x - xc + px * TWO_PI, dy = y - yc + py * TWO_PI; const double q = dx * dx + dy * dy;
            best = q < best ? q : best;
          }
          const double om = sgn * G / (M_PI * sig * sig) * exp(-best / (sig * sig));
          wz += om; wx += -sgn * a * sin(z) * om;
        }
        ox[n] = wx; oy[n] = 0; oz[n] = wz;
      }
    }
    for (int c = 0; c < 3; c++) spectral_forward(s->sp, s->om[c], s->NL[2 * c], s->NL[2 * c + 1]);
    project(s, s->NL, 1);
    curlHat(s, s->NL, s->S);
    for (int l = 0; l < N; l++) { const double kz = s->kw[l];
      for (int j = 0; j < N; j++) { const double ky = s->kw[j]; const long row = ((long)l * N + j) * NH;
        for (int i = 0; i < NH; i++) {
          const long idx = row + i; double k2 = s->kx[i] * s->kx[i] + ky * ky + kz * kz; if (k2 == 0) k2 = 1;
          for (int c = 0; c < 6; c++) s->S[c][idx] /= k2;
        } } }
    project(s, s->S, 1);
    if (kind == IC_TUBESZ && prm[3] > 0) {   /* perturbation: rms velocity prm[3] per component, seed 1, k0 = 4 */
      random_field(s, s->K, 1.0, 4.0, 1.5 * prm[3] * prm[3]);
      for (int c = 0; c < 6; c++) for (long n = 0; n < s->NS; n++) s->S[c][n] += s->K[c][n];
    }
  } else if (kind == IC_RANDOM) {
    random_field(s, s->S, prm[0], prm[1], prm[2]);
  } else return -1;
  s->st.t = 0; s->st.step = 0;
  finishIC(s);
  return 0;
}

or_solver *or_create(int N, double nu, double cfl, double dtFixed, int ic, const double *prm) {
  int f[40];
  if (or_factorize(N, f) < 0 || (N % 2)) return NULL;
  or_solver *s = (or_solver *)calloc(1, sizeof(or_solver));
  s->N = N; s->NH = N / 2 + 1; s->NS = (long)N * N * s->NH; s->N3 = (long)N * N * N;
  s->nu = nu; s->cfl = cfl; s->dtFixed = dtFixed;
  s->sp = spectral_create(N);
  s->kc = N / 3;
  s->kx = alloc0(s->NH); s->kw = alloc0(N); s->hw = alloc0(s->NH);
  for (int i = 0; i < s->NH; i++) s->kx[i] = i;
  for (int j = 0; j < N; j++) s->kw[j] = j <= N / 2 ? j : j - N;
  s->mask = (uint8_t *)calloc(s->NS, 1);
  for (int l = 0; l < N; l++) for (int j = 0; j < N; j++) for (int i = 0; i < s->NH; i++)
    s->mask[((long)l * N + j) * s->NH + i] = (fabs(s->kw[l]) <= s->kc && fabs(s->kw[j]) <= s->kc && i <= s->kc) ? 1 : 0;
  for (int i = 0; i < s->NH; i++) s->hw[i] = (i == 0 || i == N / 2) ? 1 : 2;
  for (int c = 0; c < 6; c++) { s->S[c] = alloc0(s->NS); s->T[c] = alloc0(s->NS); s->K[c] = alloc0(s->NS); s->ACC[c] = alloc0(s->NS); s->NL[c] = alloc0(s->NS); s->IW[c] = alloc0(s->NS); }
  for (int c = 0; c < 3; c++) { s->up[c] = alloc0(s->N3); s->om[c] = alloc0(s->N3); s->G[c] = alloc0(s->N3); }
  s->Sxx = alloc0(s->N3); s->Syy = alloc0(s->N3); s->Szz = alloc0(s->N3); s->Sxy = alloc0(s->N3); s->Sxz = alloc0(s->N3); s->Syz = alloc0(s->N3);
  s->G2 = alloc0(s->N3); s->PST = alloc0(s->N3);
  if (or_setIC(s, ic, prm) != 0) return NULL;
  return s;
}

void or_free(or_solver *s) {
  if (!s) return;
  spectral_free(s->sp);
  free(s->kx); free(s->kw); free(s->hw); free(s->mask);
  for (int c = 0; c < 6; c++) { free(s->S[c]); free(s->T[c]); free(s->K[c]); free(s->ACC[c]); free(s->NL[c]); free(s->IW[c]); }
  for (int c = 0; c < 3; c++) { free(s->up[c]); free(s->om[c]); free(s->G[c]); }
  free(s->Sxx); free(s->Syy); free(s->Szz); free(s->Sxy); free(s->Sxz); free(s->Syz); free(s->G2); free(s->PST);
  free(s);
}

static double chooseDt(or_solver *s) {
  if (s->dtFixed > 0) return s->dtFixed;
  const double dx = TWO_PI / s->N, uMax = s->last.uMax > 1e-9 ? s->last.uMax : 1e-9;
  const double dtc = s->cfl * dx / uMax, dtv = 2.5 / (3 * s->nu * s->kc * s->kc + 1e-30);
  return dtc < dtv ? dtc : dtv;
}

/* One RK4 step (nslab.js:312-333). Fills rec[11] = t, E, Z, eps, omMax, uMax, dt, ebal, zbal, Pspec, Pal. */
void or_step(or_solver *s, double *rec) {
  const long NS = s->NS; const double nu = s->nu;
  rhs(s, s->S, s->K);
  const or_diag d0 = s->last; const double dt = chooseDt(s); s->st.dt = dt;
  const double E0 = d0.E, Z0 = d0.Z; double eps[4], zr[4];
  eps[0] = d0.eps; zr[0] = d0.Pspec - 2 * nu * d0.P;
  { double q = fabs(d0.Tnl) / (d0.eps > 1e-300 ? d0.eps : 1e-300); if (q > s->st.maxTnl) s->st.maxTnl = q; }
  for (int c = 0; c < 6; c++) { double *S = s->S[c], *K = s->K[c], *A = s->ACC[c], *T = s->T[c]; for (long n = 0; n < NS; n++) { A[n] = K[n]; T[n] = S[n] + 0.5 * dt * K[n]; } }
  rhs(s, s->T, s->K); eps[1] = s->last.eps; zr[1] = s->last.Pspec - 2 * nu * s->last.P;
  for (int c = 0; c < 6; c++) { double *S = s->S[c], *K = s->K[c], *A = s->ACC[c], *T = s->T[c]; for (long n = 0; n < NS; n++) { A[n] += 2 * K[n]; T[n] = S[n] + 0.5 * dt * K[n]; } }
  rhs(s, s->T, s->K); eps[2] = s->last.eps; zr[2] = s->last.Pspec - 2 * nu * s->last.P;
  for (int c = 0; c < 6; c++) { double *S = s->S[c], *K = s->K[c], *A = s->ACC[c], *T = s->T[c]; for (long n = 0; n < NS; n++) { A[n] += 2 * K[n]; T[n] = S[n] + dt * K[n]; } }
  rhs(s, s->T, s->K); eps[3] = s->last.eps; zr[3] = s->last.Pspec - 2 * nu * s->last.P;
  { const double f = dt / 6; for (int c = 0; c < 6; c++) { double *S = s->S[c], *K = s->K[c], *A = s->ACC[c]; for (long n = 0; n < NS; n++) S[n] += f * (A[n] + K[n]); } }
  s->st.t += dt; s->st.step++;
  double nE, nZ, nP; norms(s, s->S, &nE, &nZ, &nP);
  const double dEpred = -dt / 6 * (eps[0] + 2 * eps[1] + 2 * eps[2] + eps[3]);
  const double dZpred = dt / 6 * (zr[0] + 2 * zr[1] + 2 * zr[2] + zr[3]);
  const double E0i = s->st.E0 > 1e-300 ? s->st.E0 : 1e-300;
  double zden = nZ > Z0 ? nZ : Z0; zden = zden > 1e-300 ? zden : 1e-300;
  const double ebal = fabs((nE - E0) - dEpred) / E0i, zbal = fabs((nZ - Z0) - dZpred) / zden;
  if (ebal > s->st.maxEbal) s->st.maxEbal = ebal;
  if (zbal > s->st.maxZbal) s->st.maxZbal = zbal;
  s->ebal = ebal; s->zbal = zbal;
  rhs(s, s->S, s->K);
  if (rec) {
    rec[0] = s->st.t; rec[1] = s->last.E; rec[2] = s->last.Z; rec[3] = s->last.eps; rec[4] = s->last.omMax; rec[5] = s->last.uMax;
    rec[6] = dt; rec[7] = ebal; rec[8] = zbal; rec[9] = s->last.Pspec; rec[10] = s->last.P;
  }
}

/* run(nSteps, tEnd) (nslab.js:334); tEnd NaN = none. Returns the number of steps taken; rec holds 11 per step. */
long or_run(or_solver *s, long nSteps, double tEnd, double *rec) {
  long k = 0;
  for (long i = 0; i < nSteps; i++) {
    if (!isnan(tEnd) && s->st.t >= tEnd - 1e-12) break;
    or_step(s, rec ? rec + 11 * k : NULL); k++;
  }
  return k;
}

/* ---------------------------------------------------------------- accessors */
void or_get_state(or_solver *s, double *out) { for (int c = 0; c < 6; c++) memcpy(out + c * s->NS, s->S[c], 8 * s->NS); }
void or_set_state(or_solver *s, const double *in) { for (int c = 0; c < 6; c++) memcpy(s->S[c], in + c * s->NS, 8 * s->NS); }
void or_get_last(or_solver *s, double *o) { o[0] = s->last.E; o[1] = s->last.Z; o[2] = s->last.P; o[3] = s->last.eps; o[4] = s->last.Tnl; o[5] = s->last.Pspec; o[6] = s->last.omMax; o[7] = s->last.uMax; }
void or_get_st(or_solver *s, double *o) { o[0] = s->st.t; o[1] = (double)s->st.step; o[2] = s->st.dt; o[3] = s->st.E0; o[4] = s->st.Z0; o[5] = s->st.maxEbal; o[6] = s->st.maxZbal; o[7] = s->st.maxTnl; o[8] = s->st.divMax; }
void or_set_st(or_solver *s, const double *o) { s->st.t = o[0]; s->st.step = (long)o[1]; s->st.dt = o[2]; s->st.E0 = o[3]; s->st.Z0 = o[4]; s->st.maxEbal = o[5]; s->st.maxZbal = o[6]; s->st.maxTnl = o[7]; s->st.divMax = o[8]; }
void or_norms(or_solver *s, double *o) { norms(s, s->S, &o[0], &o[1], &o[2]); }
/* rhs of the current state: refreshes last (as the reference's step() entry does) */
void or_rhs_current(or_solver *s) { rhs(s, s->S, s->K); }
/* physical fields after the last rhs/diagnose: 0..2 = u, v, w; 3..5 = om (u x omega after rhs, omega after diagnose) */
void or_get_physical(or_solver *s, int which, double *out) {
  const double *src = which < 3 ? s->up[which] : which < 6 ? s->om[which - 3] : which == 6 ? s->G2 : s->PST;
  memcpy(out, src, 8 * s->N3);
}

/* ---------------------------------------------------------------- diagnostics (nslab.js:339-564) */
static void eig3(double a, double b, double c, double d, double e, double f, double ev[3][3]) {
  const double p1 = d * d + e * e + f * f;
  if (p1 < 1e-30) { for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) ev[i][j] = i == j; return; }
  const double q = (a + b + c) / 3, p2 = (a - q) * (a - q) + (b - q) * (b - q) + (c - q) * (c - q) + 2 * p1, p = sqrt(p2 / 6);
  const double A = (a - q) / p, B = (b - q) / p, C = (c - q) / p, D = d / p, Ee = e / p, F = f / p;
  double r = 0.5 * (A * (B * C - F * F) - D * (D * C - F * Ee) + Ee * (D * F - B * Ee));
  r = r < -1 ? -1 : (r > 1 ? 1 : r);
  const double phi = acos(r) / 3;
  const double l1 = q + 2 * p * cos(phi), l3 = q + 2 * p * cos(phi + 2 * M_PI / 3);
  const double lams[2] = { l1, l3 };
  double vecs[2][3];
  for (int w = 0; w < 2; w++) {
    const double lam = lams[w];
    const double r0[3] = { a - lam, d, e }, r1[3] = { d, b - lam, f }, r2[3] = { e, f, c - lam };
    const double *pa[3] = { r0, r0, r1 }, *pb[3] = { r1, r2, r2 };
    double best[3] = { 0, 0, 0 }, bl = -1;
    for (int k = 0; k < 3; k++) {
      const double *x = pa[k], *y = pb[k];
      const double v0 = x[1] * y[2] - x[2] * y[1], v1 = x[2] * y[0] - x[0] * y[2], v2 = x[0] * y[1] - x[1] * y[0];
      const double l = v0 * v0 + v1 * v1 + v2 * v2;
      if (l > bl) { bl = l; best[0] = v0; best[1] = v1; best[2] = v2; }
    }
    double l = sqrt(bl); if (l == 0 || l != l) l = 1;
    vecs[w][0] = best[0] / l; vecs[w][1] = best[1] / l; vecs[w][2] = best[2] / l;
  }
  const double *e1 = vecs[0], *e3 = vecs[1];
  ev[0][0] = e1[0]; ev[0][1] = e1[1]; ev[0][2] = e1[2];
  ev[1][0] = e3[1] * e1[2] - e3[2] * e1[1]; ev[1][1] = e3[2] * e1[0] - e3[0] * e1[2]; ev[1][2] = e3[0] * e1[1] - e3[1] * e1[0];
  ev[2][0] = e3[0]; ev[2][1] = e3[1]; ev[2][2] = e3[2];
}

int or_spectrum_len(or_solver *s) { return (int)ceil(sqrt(3.0) * s->kc) + 2; }
void or_spectrum(or_solver *s, double *Ek) {
  const int N = s->N, NH = s->NH, nb = or_spectrum_len(s);
  for (int k = 0; k < nb; k++) Ek[k] = 0;
  for (int l = 0; l < N; l++) { const double kz = s->kw[l];
    for (int j = 0; j < N; j++) { const double ky = s->kw[j]; const long row = ((long)l * N + j) * NH;
      for (int i = 0; i < NH; i++) {
        const long idx = row + i; if (!s->mask[idx]) continue;
        const int k = (int)floor(sqrt(s->kx[i] * s->kx[i] + ky * ky + kz * kz) + 0.5); if (k >= nb) continue;
        double e = 0; for (int c = 0; c < 6; c++) e += s->S[c][idx] * s->S[c][idx];
        Ek[k] += 0.5 * s->hw[i] * e;
      } } }
}

/* 1 when defined (result in *out), 0 = null. kc = floor(N/3) of the grid the spectrum belongs to. */
/* The GPU runner's pileUp (nslab_gpu.py:369-372) in a working precision of machine epsilon eps: specFloor =
 * 100 eps^2 (nslab_gpu.py:64-65); nslab.js:557-564 is the eps = 2^-52 case. */
int or_cutoff_pile_up_eps(int kc, const double *spec, int nb, double eps, double *out) {
  const int k8 = (int)floor(0.8 * kc + 0.5); double peak = 0;
  for (int k = 1; k < nb; k++) peak = spec[k] > peak ? spec[k] : peak;
  const double SPEC_FLOOR = 100.0 * eps * eps;
  if (!(spec[k8] > SPEC_FLOOR * peak)) return 0;
  double m = 0; for (int k = k8; k <= kc; k++) { const double q = spec[k] / spec[k8]; m = q > m ? q : m; }
  *out = m; return 1;
}
int or_cutoff_pile_up(int kc, const double *spec, int nb, double *out) {
  return or_cutoff_pile_up_eps(kc, spec, nb, 2.220446049250313e-16, out);
}
static int cutoffPileUp(or_solver *s, const double *spec, int nb, double *out) { return or_cutoff_pile_up(s->kc, spec, nb, out); }

/* imageDiag: o[0..3] = zCentroid, zExtent, imageGap, zBands */
static void imageDiag(or_solver *s, double *o) {
  const int N = s->N; double *ox = s->om[0], *oy = s->om[1], *oz = s->om[2];
  double *Ez = alloc0(N);
  for (int l = 0; l < N; l++) { double sm = 0; for (int j = 0; j < N; j++) for (int i = 0; i < N; i++) { const long n = ((long)l * N + j) * N + i; sm += ox[n] * ox[n] + oy[n] * oy[n] + oz[n] * oz[n]; } Ez[l] = sm; }
  double mx = 0, cs = 0, sn = 0;
  for (int l = 0; l < N; l++) { mx = Ez[l] > mx ? Ez[l] : mx; const double z = l * TWO_PI / N; cs += Ez[l] * cos(z); sn += Ez[l] * sin(z); }
  const double zc = fmod(fmod(atan2(sn, cs), TWO_PI) + TWO_PI, TWO_PI);
  int *on = (int *)calloc(N, sizeof(int)); for (int l = 0; l < N; l++) on[l] = Ez[l] > 0.01 * mx;
  int (*runs)[2] = malloc(sizeof(int[2]) * (N + 1)); int nr = 0, start = -1;
  for (int l = 0; l < N; l++) { if (on[l] && start < 0) start = l; if (!on[l] && start >= 0) { runs[nr][0] = start; runs[nr][1] = l; nr++; start = -1; } }
  if (start >= 0) { runs[nr][0] = start; runs[nr][1] = N; nr++; }
  int (*rr)[2] = malloc(sizeof(int[2]) * (N + 1)); int nrr = 0;
  if (nr > 1 && on[0] && on[N - 1]) {
    rr[0][0] = runs[nr - 1][0]; rr[0][1] = runs[0][1] + N; nrr = 1;
    for (int q = 1; q < nr - 1; q++) { rr[nrr][0] = runs[q][0]; rr[nrr][1] = runs[q][1]; nrr++; }
  } else { for (int q = 0; q < nr; q++) { rr[q][0] = runs[q][0]; rr[q][1] = runs[q][1]; } nrr = nr; }
  int mxlen = 0; for (int q = 0; q < nrr; q++) { const int L = rr[q][1] - rr[q][0]; mxlen = L > mxlen ? L : mxlen; }
  double ext = (nrr ? mxlen : N) * TWO_PI / N; ext = ext < TWO_PI ? ext : TWO_PI;
  o[0] = zc; o[1] = ext; o[2] = TWO_PI - ext > 0 ? TWO_PI - ext : 0; o[3] = nrr;
  free(Ez); free(on); free(runs); free(rr);
}

/* interpMax (nslab.js:421-536): o[0..5] = omMaxI, pos x, y, z, omMaxNode */
typedef struct { double v; int i, j, l; long ord; } cand_t;
static int cand_cmp(const void *A, const void *B) {
  const cand_t *a = (const cand_t *)A, *b = (const cand_t *)B;
  if (a->v > b->v) return -1;
  if (a->v < b->v) return 1;
  return a->ord < b->ord ? -1 : (a->ord > b->ord ? 1 : 0);   /* stable: equal values keep their order */
}

typedef struct { or_solver *s; double *A, *B; } imax_ws;

static void tensors(imax_ws *ws, const double x[3], int c, double T[27]) {
  or_solver *s = ws->s; const int N = s->N, NH = s->NH;
  const double *wr = s->IW[2 * c], *wi = s->IW[2 * c + 1];
  double *A = ws->A, *B = ws->B;
  memset(A, 0, sizeof(double) * 2 * N * N * 3);
  for (int l = 0; l < N; l++) for (int j = 0; j < N; j++) {
    const long row = ((long)l * N + j) * NH; double s0r = 0, s0i = 0, s1r = 0, s1i = 0, s2r = 0, s2i = 0;
    for (int i = 0; i < NH; i++) {
      const long idx = row + i; const double kxx = s->kx[i], ph = kxx * x[0], cr = cos(ph), ci = sin(ph), w = s->hw[i];
      const double ar = w * (wr[idx] * cr - wi[idx] * ci), ai = w * (wr[idx] * ci + wi[idx] * cr);
      s0r += ar; s0i += ai;
      s1r += -kxx * ai; s1i += kxx * ar;
      s2r += -kxx * kxx * ar; s2i += -kxx * kxx * ai;
    }
    const long b = (((long)l * N + j) * 3) * 2;
    A[b] = s0r; A[b + 1] = s0i; A[b + 2] = s1r; A[b + 3] = s1i; A[b + 4] = s2r; A[b + 5] = s2i;
  }
  memset(B, 0, sizeof(double) * 2 * N * 3 * 3);
  for (int l = 0; l < N; l++)
    for (int j = 0; j < N; j++) {
      const double ky = s->kw[j], ph = ky * x[1], cy = cos(ph), sy = sin(ph); const long base = (((long)l * N + j) * 3) * 2;
      for (int a = 0; a < 3; a++) {
        const double ar0 = A[base + 2 * a], ai0 = A[base + 2 * a + 1];
        const double er = ar0 * cy - ai0 * sy, ei = ar0 * sy + ai0 * cy;
        for (int bb = 0; bb + a < 3; bb++) {
          const double f = bb == 0 ? 1 : bb == 1 ? ky : -ky * ky; const int rot = bb == 1;
          const double rr = rot ? -f * ei : f * er, ri = rot ? f * er : f * ei;
          const long o = (((long)l * 3 + a) * 3 + bb) * 2; B[o] += rr; B[o + 1] += ri;
        }
      }
    }
  for (int q = 0; q < 27; q++) T[q] = 0;
  for (int l = 0; l < N; l++) {
    const double kz = s->kw[l], ph = kz * x[2], cz = cos(ph), sz = sin(ph);
    for (int a = 0; a < 3; a++) for (int bb = 0; bb + a < 3; bb++) {
      const long o = (((long)l * 3 + a) * 3 + bb) * 2; const double br0 = B[o], bi0 = B[o + 1];
      const double er = br0 * cz - bi0 * sz, ei = br0 * sz + bi0 * cz;
      for (int cc = 0; cc + a + bb < 3; cc++) {
        const double f = cc == 0 ? 1 : cc == 1 ? kz : -kz * kz; const int rot = cc == 1;
        T[(a * 3 + bb) * 3 + cc] += rot ? -f * ei : f * er;
      }
    }
  }
}

typedef struct { double f, g[3], H[3][3]; } fgh_t;
static fgh_t fgh(imax_ws *ws, const double x[3]) {
  fgh_t r; memset(&r, 0, sizeof r);
  for (int c = 0; c < 3; c++) {
    double T[27]; tensors(ws, x, c, T);
    const double w = T[0];
    const double gw[3] = { T[(1 * 3 + 0) * 3 + 0], T[(0 * 3 + 1) * 3 + 0], T[(0 * 3 + 0) * 3 + 1] };
    const double Hw[3][3] = { { T[(2 * 3 + 0) * 3 + 0], T[(1 * 3 + 1) * 3 + 0], T[(1 * 3 + 0) * 3 + 1] },
                              { T[(1 * 3 + 1) * 3 + 0], T[(0 * 3 + 2) * 3 + 0], T[(0 * 3 + 1) * 3 + 1] },
                              { T[(1 * 3 + 0) * 3 + 1], T[(0 * 3 + 1) * 3 + 1], T[(0 * 3 + 0) * 3 + 2] } };
    r.f += w * w;
    for (int p = 0; p < 3; p++) { r.g[p] += 2 * w * gw[p]; for (int q = 0; q < 3; q++) r.H[p][q] += 2 * (gw[p] * gw[q] + w * Hw[p][q]); }
  }
  return r;
}
static double det3(double m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
static int solve3(double M[3][3], const double r[3], double out[3]) {
  const double d = det3(M);
  if (fabs(d) < 1e-300) return 0;
  for (int i = 0; i < 3; i++) {
    double A2[3][3]; memcpy(A2, M, sizeof A2);
    for (int q = 0; q < 3; q++) A2[q][i] = r[q];
    out[i] = det3(A2) / d;
  }
  return 1;
}

void or_interpMax(or_solver *s, int top, int iters, double *o) {
  const int N = s->N; const long N3 = s->N3;
  curlHat(s, s->S, s->IW);
  double *ox = s->om[0], *oy = s->om[1], *oz = s->om[2];
#define O2(n) (ox[n] * ox[n] + oy[n] * oy[n] + oz[n] * oz[n])
  cand_t *cand = (cand_t *)malloc(sizeof(cand_t) * N3), *flat = (cand_t *)malloc(sizeof(cand_t) * N3);
  long nc = 0;
  for (int l = 0; l < N; l++) for (int j = 0; j < N; j++) for (int i = 0; i < N; i++) {
    const long n = ((long)l * N + j) * N + i; const double v = O2(n);
    const int lm = v >= O2((((long)(l + 1) % N) * N + j) * N + i) && v >= O2((((long)(l + N - 1) % N) * N + j) * N + i)
      && v >= O2(((long)l * N + (j + 1) % N) * N + i) && v >= O2(((long)l * N + (j + N - 1) % N) * N + i)
      && v >= O2(((long)l * N + j) * N + (i + 1) % N) && v >= O2(((long)l * N + j) * N + (i + N - 1) % N);
    if (lm) { cand[nc].v = v; cand[nc].i = i; cand[nc].j = j; cand[nc].l = l; cand[nc].ord = nc; nc++; }
    flat[n].v = v; flat[n].i = i; flat[n].j = j; flat[n].l = l; flat[n].ord = n;
  }
  qsort(flat, N3, sizeof(cand_t), cand_cmp); qsort(cand, nc, sizeof(cand_t), cand_cmp);
  cand_t *starts = (cand_t *)malloc(sizeof(cand_t) * 2 * top); int ns = 0;
  const int t1 = nc < top ? (int)nc : top, t2 = N3 < top ? (int)N3 : top;
  for (int q = 0; q < t1 + t2; q++) {
    const cand_t c = q < t1 ? cand[q] : flat[q - t1];
    int dup = 0; for (int r = 0; r < ns; r++) if (starts[r].i == c.i && starts[r].j == c.j && starts[r].l == c.l) { dup = 1; break; }
    if (!dup) { starts[ns] = c; starts[ns].ord = ns; ns++; }
  }
  qsort(starts, ns, sizeof(cand_t), cand_cmp);
  const double h = TWO_PI / N;
  imax_ws ws = { s, alloc0(2L * N * N * 3), alloc0(2L * N * 3 * 3) };
  int have = 0; double bf = 0, bx[3] = { 0, 0, 0 }, bnode = 0;
  for (int q = 0; q < ns; q++) {
    const double v0 = starts[q].v;
    if (have && sqrt(v0) < 0.85 * sqrt(bf)) continue;
    double x[3] = { starts[q].i * h, starts[q].j * h, starts[q].l * h };
    fgh_t cur = fgh(&ws, x);
    for (int it = 0; it < iters; it++) {
      double stp[3]; const double mg[3] = { -cur.g[0], -cur.g[1], -cur.g[2] };
      const int ok = solve3(cur.H, mg, stp);
      const double gn = sqrt(cur.g[0] * cur.g[0] + cur.g[1] * cur.g[1] + cur.g[2] * cur.g[2]);
      if (!ok || stp[0] * cur.g[0] + stp[1] * cur.g[1] + stp[2] * cur.g[2] <= 0) {
        const double den = gn > 1e-300 ? gn : 1e-300;
        for (int p = 0; p < 3; p++) stp[p] = cur.g[p] / den * 0.25 * h;
      }
      double n = sqrt(stp[0] * stp[0] + stp[1] * stp[1] + stp[2] * stp[2]);
      if (n > h) { for (int p = 0; p < 3; p++) stp[p] = stp[p] * h / n; n = h; }
      double xt[3] = { x[0] + stp[0], x[1] + stp[1], x[2] + stp[2] };
      fgh_t trial = fgh(&ws, xt);
      while (trial.f < cur.f && n > 1e-9 * h) {
        for (int p = 0; p < 3; p++) stp[p] *= 0.5;
        n *= 0.5;
        for (int p = 0; p < 3; p++) xt[p] = x[p] + stp[p];
        trial = fgh(&ws, xt);
      }
      if (trial.f < cur.f) break;
      for (int p = 0; p < 3; p++) x[p] = x[p] + stp[p];
      cur = trial;
      if (n < 1e-7 * h) break;
    }
    if (!have || cur.f > bf) { have = 1; bf = cur.f; bx[0] = x[0]; bx[1] = x[1]; bx[2] = x[2]; bnode = v0; }
  }
  o[0] = sqrt(bf > 0 ? bf : 0);
  for (int p = 0; p < 3; p++) o[1 + p] = fmod(fmod(bx[p], TWO_PI) + TWO_PI, TWO_PI);
  o[4] = sqrt(bnode > 0 ? bnode : 0);
  free(cand); free(flat); free(starts); free(ws.A); free(ws.B);
#undef O2
}

/* diagnose (nslab.js:339-392). out layout (doubles):
 *  0 t, 1 step, 2 E, 3 Z, 4 P, 5 eps, 6 omMax, 7 uMax, 8 Pspec, 9 Pphys, 10 Tnl, 11 skew, 12 divMax, 13 eta,
 *  14 kmaxEta, 15 Rel, 16 lambda, 17 skewIso, 18..20 align, 21 uL3, 22 pileUp (NaN = null), 23 zCentroid,
 *  24 zExtent, 25 imageGap, 26 zBands, 27 omMaxI (NaN if not run), 28..30 omMaxIpos, 31 omMaxNode,
 *  32..79 alignHist[3][16], 80.. spectrum[nb] */
void or_diagnose(or_solver *s, int doInterp, double *out) {
  const int N = s->N, NH = s->NH; const long N3 = s->N3; const double nu = s->nu;
  rhs(s, s->S, s->K);
  curlHat(s, s->S, s->NL);
  for (int c = 0; c < 3; c++) spectral_inverse(s->sp, s->NL[2 * c], s->NL[2 * c + 1], s->om[c]);
  double *u = s->up[0], *v = s->up[1], *w = s->up[2], *ox = s->om[0], *oy = s->om[1], *oz = s->om[2];
  memset(s->Sxx, 0, 8 * N3); memset(s->Syy, 0, 8 * N3); memset(s->Szz, 0, 8 * N3); memset(s->Sxy, 0, 8 * N3);
  memset(s->Sxz, 0, 8 * N3); memset(s->Syz, 0, 8 * N3); memset(s->G2, 0, 8 * N3); memset(s->PST, 0, 8 * N3);
  double skew3 = 0, skew2 = 0, divMax = 0;
  double *tr = s->NL[0], *ti = s->NL[1];
  for (int c = 0; c < 3; c++) {
    const double *ur = s->S[2 * c], *ui = s->S[2 * c + 1];
    for (int jd = 0; jd < 3; jd++) {
      for (int l = 0; l < N; l++) { const double kz = s->kw[l];
        for (int j = 0; j < N; j++) { const double ky = s->kw[j]; const long row = ((long)l * N + j) * NH;
          for (int i = 0; i < NH; i++) { const long idx = row + i; const double kk = jd == 0 ? s->kx[i] : jd == 1 ? ky : kz; tr[idx] = -kk * ui[idx]; ti[idx] = kk * ur[idx]; } } }
      spectral_inverse(s->sp, tr, ti, s->G[jd]);
    }
    const double *oc = s->om[c], *gx = s->G[0], *gy = s->G[1], *gz = s->G[2];
    for (long n = 0; n < N3; n++) {
      const double a = gx[n], b = gy[n], d = gz[n];
      s->G2[n] += a * a + b * b + d * d;
      s->PST[n] += oc[n] * (ox[n] * a + oy[n] * b + oz[n] * d);
      const double dg = c == 0 ? a : c == 1 ? b : d; skew3 += dg * dg * dg; skew2 += dg * dg;
    }
    if (c == 0) { for (long n = 0; n < N3; n++) { s->Sxx[n] = gx[n]; s->Sxy[n] += 0.5 * gy[n]; s->Sxz[n] += 0.5 * gz[n]; } }
    else if (c == 1) { for (long n = 0; n < N3; n++) { s->Syy[n] = gy[n]; s->Sxy[n] += 0.5 * gx[n]; s->Syz[n] += 0.5 * gz[n]; } }
    else { for (long n = 0; n < N3; n++) { s->Szz[n] = gz[n]; s->Sxz[n] += 0.5 * gx[n]; s->Syz[n] += 0.5 * gy[n]; } }
  }
  double Pphys = 0, align[3] = { 0, 0, 0 }, cnt = 0, u3 = 0, hist[3][16];
  memset(hist, 0, sizeof hist);
  for (long n = 0; n < N3; n++) {
    const double dv = fabs(s->Sxx[n] + s->Syy[n] + s->Szz[n]); if (dv > divMax) divMax = dv;
    const double o2 = ox[n] * ox[n] + oy[n] * oy[n] + oz[n] * oz[n];
    u3 += pow(u[n] * u[n] + v[n] * v[n] + w[n] * w[n], 1.5);
    s->G2[n] = 0.25 * o2 - 0.5 * s->G2[n];
    Pphys += s->PST[n];
    if (o2 > 1e-20) {
      double ev[3][3]; eig3(s->Sxx[n], s->Syy[n], s->Szz[n], s->Sxy[n], s->Sxz[n], s->Syz[n], ev);
      const double io = 1 / sqrt(o2);
      for (int e = 0; e < 3; e++) {
        const double cs = fabs((ox[n] * ev[e][0] + oy[n] * ev[e][1] + oz[n] * ev[e][2]) * io);
        align[e] += cs; int bin = (int)floor(cs * 16); bin = bin < 15 ? bin : 15; hist[e][bin] += 1;
      }
      cnt++;
    }
  }
  if (divMax > s->st.divMax) s->st.divMax = divMax;
  const int nb = or_spectrum_len(s);
  double *spec = out + 80;
  or_spectrum(s, spec);
  const or_diag d = s->last;
  const double epsc = d.eps > 1e-300 ? d.eps : 1e-300;
  const double eta = pow(nu * nu * nu / epsc, 0.25), kmax = sqrt(3.0) * s->kc;
  const double lam = sqrt(10 * nu * d.E / epsc), Rel = sqrt(2 * d.E / 3) * lam / nu;
  const double cn = cnt > 1 ? cnt : 1;
  out[0] = s->st.t; out[1] = (double)s->st.step; out[2] = d.E; out[3] = d.Z; out[4] = d.P; out[5] = d.eps; out[6] = d.omMax; out[7] = d.uMax;
  out[8] = d.Pspec; out[9] = Pphys / N3; out[10] = d.Tnl;
  out[11] = skew2 > 0 ? (skew3 / 3) / pow(skew2 / 3, 1.5) : 0;
  out[12] = divMax; out[13] = eta; out[14] = kmax * eta; out[15] = Rel; out[16] = lam;
  { const double z2 = 2 * d.Z > 1e-300 ? 2 * d.Z : 1e-300; out[17] = -(6 * sqrt(15.0) / 7) * d.Pspec / pow(z2, 1.5); }
  for (int e = 0; e < 3; e++) out[18 + e] = align[e] / cn;
  out[21] = pow(u3 / N3, 1.0 / 3);
  { double pu; out[22] = cutoffPileUp(s, spec, nb, &pu) ? pu : NAN; }
  imageDiag(s, out + 23);
  if (doInterp && N <= 128) or_interpMax(s, 24, 12, out + 27);
  else { out[27] = NAN; out[28] = out[29] = out[30] = NAN; out[31] = NAN; }
  for (int e = 0; e < 3; e++) for (int b = 0; b < 16; b++) out[32 + e * 16 + b] = hist[e][b] / cn;
}

/* exactError (nslab.js:585-594): o[0] = linf, o[1] = rel; returns 0 if the IC has no exact solution */
int or_exactError(or_solver *s, double *o) {
  const int N = s->N; const double h = TWO_PI / N;
  for (int c = 0; c < 3; c++) spectral_inverse(s->sp, s->S[2 * c], s->S[2 * c + 1], s->up[c]);
  double *u = s->up[0], *v = s->up[1], *w = s->up[2], err = 0, ref = 0;
#define MX(a, b) ((a) > (b) ? (a) : (b))
  if (s->ic == IC_TGV2D) {
    const double f = exp(-2 * s->nu * s->st.t);
    for (int l = 0; l < N; l++) for (int j = 0; j < N; j++) for (int i = 0; i < N; i++) {
      const long n = ((long)l * N + j) * N + i; const double x = i * h, y = j * h;
      const double ue = sin(x) * cos(y) * f, ve = -cos(x) * sin(y) * f;
      err = MX(err, MX(fabs(u[n] - ue), MX(fabs(v[n] - ve), fabs(w[n])))); ref = MX(ref, fabs(ue));
    }
  } else if (s->ic == IC_ABC) {
    const double A = s->icp[0], B = s->icp[1], C = s->icp[2], f = exp(-s->nu * s->st.t);
    for (int l = 0; l < N; l++) for (int j = 0; j < N; j++) for (int i = 0; i < N; i++) {
      const long n = ((long)l * N + j) * N + i; const double x = i * h, y = j * h, z = l * h;
      const double ue = (A * sin(z) + C * cos(y)) * f, ve = (B * sin(x) + A * cos(z)) * f, we = (C * sin(y) + B * cos(x)) * f;
      err = MX(err, MX(fabs(u[n] - ue), MX(fabs(v[n] - ve), fabs(w[n] - we)))); ref = MX(ref, MX(fabs(ue), MX(fabs(ve), fabs(we))));
    }
  } else return 0;
  o[0] = err; o[1] = err / MX(ref, 1e-300);
#undef MX
  return 1;
}

int or_N(or_solver *s) { return s->N; }
long or_NS(or_solver *s) { return s->NS; }
