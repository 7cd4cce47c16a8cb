// Kernels of the NSLab gfx950 solver. One right-hand-side evaluation (nslab.js:224-247) is five pencil passes plus
// one fused spectral sweep, each a single HBM read and write of its operands:
//   P1 k_zinv    z-inverse of the compact state, curl formed on the fly        3 fields -> 6 (u, omega)
//   P2 k_yinv    y-inverse                                                      6 -> 6
//   P3 k_xnl     x C2R of u and omega, max|u|^2, max|omega|^2, u x omega, x R2C  6 -> 3 (in place)
//   P4 k_yfwd    y-forward, truncated to the retained ky                        3 -> 3
//   P5 k_zfwd    z-forward, truncated to the retained kz, times 1/N^3          3 -> NL (compact)
//   k_update<s>  Leray projection (+ the 2/3 mask, implicit in compact storage), viscous term, RK4 stage s
//                arithmetic, and the Tnl/Pspec/E/Z/Pal (and new-state E/Z/P) partial sums
// The step control (dt from CFL, t, step, budgets, stop predicate) lives on the device in Ctl so a whole run is
// enqueued without host round trips.
#pragma once
#include "nsl_device.hpp"

namespace nsl {

#define NSL_TWO_PI 6.283185307179586

// One per solver, in device memory.
struct Ctl {
  int active;            // the step in flight executes (0: every kernel of the step returns at once)
  int pend;              // the fused update of the last rhs left its partial sums unbooked (see fin_rhs)
  long long nDone;       // steps done in the current run call
  long long nTarget;     // steps requested
  long long nChunk;      // records written since the last chunk start
  double tEnd;
  int hasTEnd;
  int N;
  int kc;
  int pad;
  double nu, cfl, dtFixed;
  double t, dt;
  long long step;
  double E0ref;
  unsigned long long om2, u2;   // max |omega|^2, |u|^2 of the rhs in flight (IEEE bits)
  double diag[5][8];            // stages 1..4 of the step in flight, [4] = `last`: E Z P eps Tnl Pspec omMax uMax
  double nrm[3];                // E Z P of the state after the RK4 update
  double ebal, zbal;
};

struct StepRecDev {
  double t, dt;
  long long step;
  double ebal, zbal;
  double last[8];
  double d0[8];
};

// Storage element of the spectral arrays: double2 (FP64 solves) or float2 (FP32 solves: complex64 in HBM; the update
// and the sums work in FP64 registers, the transforms in float2). ldc widens a stored value, stc rounds a value to the storage precision, and
// rnd<E> is the value a store to E followed by a load gives back (the identity for double2).
__device__ __forceinline__ double2 ldc(const double2 &v) { return v; }
__device__ __forceinline__ double2 ldc(const float2 &v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ void stc(double2 &d, const double2 &v) { d = v; }
__device__ __forceinline__ void stc(float2 &d, const double2 &v) { d = make_float2((float)v.x, (float)v.y); }
template <class E> __device__ __forceinline__ double2 rnd(const double2 &v) {
  if constexpr (std::is_same<E, double2>::value) return v;
  else return make_double2((double)(float)v.x, (double)(float)v.y);
}
// The same between a stored element E and a compute value C of any precision (the pass kernels of FP32 solves
// transform in float2): ldx reads a stored value as C, stx stores a C value as E (rounded when E is narrower).
template <class C, class E> __device__ __forceinline__ C ldx(const E &v) {
  if constexpr (std::is_same<C, E>::value) return v;
  else { C r; r.x = v.x; r.y = v.y; return r; }
}
template <class E, class C> __device__ __forceinline__ void stx(E &d, const C &v) {
  if constexpr (std::is_same<C, E>::value) d = v;
  else { E r; r.x = v.x; r.y = v.y; d = r; }
}

// chooseDt (nslab.js:305-310), evaluated with the reference's operation order.
__device__ __forceinline__ double choose_dt(const Ctl *c, double uMaxIn) {
  if (c->dtFixed > 0) return c->dtFixed;
  const double dx = NSL_TWO_PI / c->N, uMax = fmax(uMaxIn, 1e-9);
  const double dtc = c->cfl * dx / uMax, dtv = 2.5 / (3 * c->nu * c->kc * c->kc + 1e-30);
  return fmin(dtc, dtv);
}

// ------------------------------------------------------------------------------------------------ P1
// Source of output component o at compact mode m. MODE 0: field o as stored. MODE 1: o < 3 velocity, o >= 3 the
// vorticity i k x u (curlHat, nslab.js:203-212). MODE 2: the gradient i k_o u_gc (diagnose, nslab.js:352-354).
template <int MODE>
__device__ __forceinline__ double2 zsrc(const double2 *__restrict__ U, long long D, size_t m, int o, double kx,
                                        double ky, double kz, int gc) {
  if (MODE == 0) return U[(size_t)o * D + m];
  if (MODE == 2) {
    const double2 u = U[(size_t)gc * D + m];
    const double kk = o == 0 ? kx : (o == 1 ? ky : kz);
    return make_double2(-kk * u.y, kk * u.x);
  }
  if (o < 3) return U[(size_t)o * D + m];
  const double2 a = U[m], b = U[(size_t)D + m], c = U[2 * (size_t)D + m];
  double pr, pi;
  if (o == 3) { pr = ky * c.x - kz * b.x; pi = ky * c.y - kz * b.y; }
  else if (o == 4) { pr = kz * a.x - kx * c.x; pi = kz * a.y - kx * c.y; }
  else { pr = kx * b.x - ky * a.x; pi = kx * b.y - ky * a.y; }
  return make_double2(-pi, pr);
}

// Grid: nout * ntile * by blocks (component fastest so the tiles that share inputs run together).
template <int TN, int MODE>
__global__ __launch_bounds__(NSL_NT) void k_zinv(const double2 *__restrict__ U, double2 *__restrict__ X,
                                                  const Geo g, const DPlan P, int lgT, int nout, int gc,
                                                  const Ctl *__restrict__ ctl) {
  if (ctl != nullptr && ctl->active == 0) return;
  extern __shared__ double2 lds[];
  const int T = 1 << lgT, N = g.N;
  const int ntile = (g.hx + T - 1) >> lgT;
  const int o = blockIdx.x % nout;
  const int rest = blockIdx.x / nout;
  const int tile = rest % ntile, cy = rest / ntile;
  const int kx0 = tile << lgT;
  const double ky = kwv(full_of(cy, g), N);
  for (int e = threadIdx.x; e < (N << lgT); e += NSL_NT) {
    const int n = e >> lgT, t = e & (T - 1), kx = kx0 + t;
    double2 v = make_double2(0.0, 0.0);
    const int cz = compact_of(n, g);
    if (cz >= 0 && kx < g.hx) {
      const size_t m = ((size_t)cz * g.by + cy) * g.hp + kx;
      v = zsrc<MODE>(U, g.D, m, o, (double)kx, ky, kwv(n, N), gc);
    }
    lds[e] = v;
  }
  __syncthreads();
  lds_fft<TN>(lds, T, P, 1.0);
  for (int e = threadIdx.x; e < (N << lgT); e += NSL_NT) {
    const int n = e >> lgT, t = e & (T - 1), kx = kx0 + t;
    if (kx < g.hx) X[(((size_t)o * N + n) * g.by + cy) * g.hp + kx] = lds[e];
  }
}

// ------------------------------------------------------------------------------------------------ P2
// X1[o][z][cy][kx] -> X2[o][z][y][kx]. Grid: (ntile*N, ncomp).
template <int TN>
__global__ __launch_bounds__(NSL_NT) void k_yinv(const double2 *__restrict__ X1, double2 *__restrict__ X2,
                                                  const Geo g, const DPlan P, int lgT, const Ctl *__restrict__ ctl) {
  if (ctl != nullptr && ctl->active == 0) return;
  extern __shared__ double2 lds[];
  const int T = 1 << lgT, N = g.N;
  const int ntile = (g.hx + T - 1) >> lgT;
  const int tile = blockIdx.x % ntile, z = blockIdx.x / ntile, o = blockIdx.y;
  const int kx0 = tile << lgT;
  const size_t inb = ((size_t)o * N + z) * g.by, outb = ((size_t)o * N + z) * N;
  for (int e = threadIdx.x; e < (N << lgT); e += NSL_NT) {
    const int n = e >> lgT, t = e & (T - 1), kx = kx0 + t;
    const int cy = compact_of(n, g);
    lds[e] = (cy >= 0 && kx < g.hx) ? X1[(inb + cy) * g.hp + kx] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  lds_fft<TN>(lds, T, P, 1.0);
  for (int e = threadIdx.x; e < (N << lgT); e += NSL_NT) {
    const int n = e >> lgT, t = e & (T - 1), kx = kx0 + t;
    if (kx < g.hx) X2[(outb + n) * g.hp + kx] = lds[e];
  }
}

// ------------------------------------------------------------------------------------------------ P3 helpers
// Hermitian rebuild of two half-spectra A (row y) and B (row y+1) into one complex line Z = A + iB (nslab.js:144-150):
// Z[k] = A + iB, Z[N-k] = conj(A) + i conj(B). LDS must be zero beforehand (modes hx..N-hx are zero).
__device__ __forceinline__ void pack_pair(double2 *lds, int T, int t, int k, int N, double2 A, double2 B) {
  lds[k * T + t] = make_double2(A.x - B.y, A.y + B.x);
  if (k > 0 && 2 * k < N) lds[(N - k) * T + t] = make_double2(A.x + B.y, B.x - A.y);
}
// Split the forward transform Z of f(y) + i f(y+1) into the two half-spectra at k (nslab.js:110-113).
__device__ __forceinline__ void unpack_pair(const double2 *lds, int T, int t, int k, int N, double2 &X, double2 &Y) {
  const double2 Zk = lds[k * T + t], Zc = lds[((N - k) % N) * T + t];
  const double zr = Zk.x, zi = Zk.y, cr = Zc.x, ci = -Zc.y;
  X = make_double2(0.5 * (zr + cr), 0.5 * (zi + ci));
  Y = make_double2(0.5 * (zi - ci), -0.5 * (zr - cr));
}

// P3: X2[6][N][N][hx] (u, omega after the y and z inverses) -> X2[0..2] = x-R2C of u x omega. One workgroup takes
// LP pairs of rows (z, y), (z, y+1); the two rows ride as real and imaginary part of one complex transform, as the
// reference packs them (nslab.js:104-114, 144-154). Grid: ceil(N*N/2 / LP).
template <int TN>
__global__ __launch_bounds__(NSL_NT) void k_xnl(double2 *__restrict__ X2, const Geo g, const DPlan P, int LP,
                                                 Ctl *__restrict__ ctl) {
  if (ctl->active == 0) return;
  extern __shared__ double2 lds[];
  const int N = g.N, hx = g.hx, half = N >> 1, T = 6 * LP, T2 = 3 * LP;
  const int npairs = N * half;
  const size_t cs = (size_t)N * N * g.hp;   // component stride
  for (int e = threadIdx.x; e < N * T; e += NSL_NT) lds[e] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int e = threadIdx.x; e < T * hx; e += NSL_NT) {
    const int t = e / hx, k = e - t * hx, p = t / 6, o = t - p * 6;
    const int pr = blockIdx.x * LP + p;
    if (pr >= npairs) continue;
    const int z = pr / half, y = 2 * (pr - z * half);
    const size_t base = o * cs + ((size_t)z * N + y) * g.hp;
    pack_pair(lds, T, t, k, N, X2[base + k], X2[base + g.hp + k]);
  }
  __syncthreads();
  lds_fft<TN>(lds, T, P, 1.0);
  constexpr int KI = (TN / 6 + NSL_NT - 1) / NSL_NT;
  double nv[KI][6];
  double om2 = 0.0, u2 = 0.0;
#pragma unroll
  for (int k2 = 0; k2 < KI; k2++) {
    const int it = threadIdx.x + k2 * NSL_NT;
    if (it < N * LP) {
      const int x = it / LP, p = it - x * LP;
      const double2 *Z = lds + x * T + p * 6;
      const double2 Za = Z[0], Zb = Z[1], Zc = Z[2], Zp = Z[3], Zq = Z[4], Zr = Z[5];
      {  // row y: real parts (nslab.js:229-233)
        const double a = Za.x, b = Zb.x, c = Zc.x, pp = Zp.x, q = Zq.x, r = Zr.x;
        const double o2 = pp * pp + q * q + r * r, uu = a * a + b * b + c * c;
        om2 = fmax(om2, o2); u2 = fmax(u2, uu);
        nv[k2][0] = b * r - c * q; nv[k2][1] = c * pp - a * r; nv[k2][2] = a * q - b * pp;
      }
      {  // row y+1: imaginary parts
        const double a = Za.y, b = Zb.y, c = Zc.y, pp = Zp.y, q = Zq.y, r = Zr.y;
        const double o2 = pp * pp + q * q + r * r, uu = a * a + b * b + c * c;
        om2 = fmax(om2, o2); u2 = fmax(u2, uu);
        nv[k2][3] = b * r - c * q; nv[k2][4] = c * pp - a * r; nv[k2][5] = a * q - b * pp;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < KI; k2++) {
    const int it = threadIdx.x + k2 * NSL_NT;
    if (it < N * LP) {
      const int x = it / LP, p = it - x * LP;
      double2 *W = lds + x * T2 + p * 3;
      W[0] = make_double2(nv[k2][0], nv[k2][3]);
      W[1] = make_double2(nv[k2][1], nv[k2][4]);
      W[2] = make_double2(nv[k2][2], nv[k2][5]);
    }
  }
  __syncthreads();
  lds_fft<TN>(lds, T2, P, -1.0);
  for (int e = threadIdx.x; e < T2 * hx; e += NSL_NT) {
    const int t = e / hx, k = e - t * hx, p = t / 3, c = t - p * 3;
    const int pr = blockIdx.x * LP + p;
    if (pr >= npairs) continue;
    const int z = pr / half, y = 2 * (pr - z * half);
    double2 X, Y;
    unpack_pair(lds, T2, t, k, N, X, Y);
    const size_t base = c * cs + ((size_t)z * N + y) * g.hp;
    X2[base + k] = X;
    X2[base + g.hp + k] = Y;
  }
  block_atomic_max2(om2, u2, &ctl->om2, &ctl->u2, (double *)lds);
}

// x R2C of ncomp real fields phys[c][z][y][x] -> X[c][z][y][0..hx). Grid: ceil(N*N/2 / LP).
template <int TN>
__global__ __launch_bounds__(NSL_NT) void k_xfwd(const double *__restrict__ phys, double2 *__restrict__ X,
                                                  const Geo g, const DPlan P, int LP, int ncomp) {
  extern __shared__ double2 lds[];
  const int N = g.N, hx = g.hx, half = N >> 1, T = ncomp * LP;
  const int npairs = N * half;
  const size_t ps = (size_t)N * N * N, cs = (size_t)N * N * g.hp;
  for (int e = threadIdx.x; e < N * T; e += NSL_NT) {
    const int x = e / T, t = e - x * T, p = t / ncomp, c = t - p * ncomp;
    const int pr = blockIdx.x * LP + p;
    double2 v = make_double2(0.0, 0.0);
    if (pr < npairs) {
      const int z = pr / half, y = 2 * (pr - z * half);
      const size_t b0 = c * ps + ((size_t)z * N + y) * N;
      v = make_double2(phys[b0 + x], phys[b0 + N + x]);
    }
    lds[e] = v;
  }
  __syncthreads();
  lds_fft<TN>(lds, T, P, -1.0);
  for (int e = threadIdx.x; e < T * hx; e += NSL_NT) {
    const int t = e / hx, k = e - t * hx, p = t / ncomp, c = t - p * ncomp;
    const int pr = blockIdx.x * LP + p;
    if (pr >= npairs) continue;
    const int z = pr / half, y = 2 * (pr - z * half);
    double2 A, B;
    unpack_pair(lds, T, t, k, N, A, B);
    const size_t base = c * cs + ((size_t)z * N + y) * g.hp;
    X[base + k] = A;
    X[base + g.hp + k] = B;
  }
}

// x C2R: X[c][z][y][0..hx) -> phys[c][z][y][x] for ncomp fields. Grid: ceil(N*N/2 / LP).
template <int TN>
__global__ __launch_bounds__(NSL_NT) void k_xinv(const double2 *__restrict__ X, double *__restrict__ phys,
                                                  const Geo g, const DPlan P, int LP, int ncomp) {
  extern __shared__ double2 lds[];
  const int N = g.N, hx = g.hx, half = N >> 1, T = ncomp * LP;
  const int npairs = N * half;
  const size_t ps = (size_t)N * N * N, cs = (size_t)N * N * g.hp;
  for (int e = threadIdx.x; e < N * T; e += NSL_NT) lds[e] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int e = threadIdx.x; e < T * hx; e += NSL_NT) {
    const int t = e / hx, k = e - t * hx, p = t / ncomp, c = t - p * ncomp;
    const int pr = blockIdx.x * LP + p;
    if (pr >= npairs) continue;
    const int z = pr / half, y = 2 * (pr - z * half);
    const size_t base = c * cs + ((size_t)z * N + y) * g.hp;
    pack_pair(lds, T, t, k, N, X[base + k], X[base + g.hp + k]);
  }
  __syncthreads();
  lds_fft<TN>(lds, T, P, 1.0);
  for (int e = threadIdx.x; e < N * T; e += NSL_NT) {
    const int x = e / T, t = e - x * T, p = t / ncomp, c = t - p * ncomp;
    const int pr = blockIdx.x * LP + p;
    if (pr >= npairs) continue;
    const int z = pr / half, y = 2 * (pr - z * half);
    const size_t b0 = c * ps + ((size_t)z * N + y) * N;
    const double2 v = lds[e];
    phys[b0 + x] = v.x;
    phys[b0 + N + x] = v.y;
  }
}

// ------------------------------------------------------------------------------------------------ P4, P5
// X3[c][z][y][kx] -> X4[c][z][cy][kx] (retained ky only). Grid: (ntile*N, ncomp).
template <int TN>
__global__ __launch_bounds__(NSL_NT) void k_yfwd(const double2 *__restrict__ X3, double2 *__restrict__ X4,
                                                  const Geo g, const DPlan P, int lgT, const Ctl *__restrict__ ctl) {
  if (ctl != nullptr && ctl->active == 0) return;
  extern __shared__ double2 lds[];
  const int T = 1 << lgT, N = g.N;
  const int ntile = (g.hx + T - 1) >> lgT;
  const int tile = blockIdx.x % ntile, z = blockIdx.x / ntile, c = blockIdx.y;
  const int kx0 = tile << lgT;
  const size_t inb = ((size_t)c * N + z) * N, outb = ((size_t)c * N + z) * g.by;
  for (int e = threadIdx.x; e < (N << lgT); e += NSL_NT) {
    const int n = e >> lgT, t = e & (T - 1), kx = kx0 + t;
    lds[e] = kx < g.hx ? X3[(inb + n) * g.hp + kx] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  lds_fft<TN>(lds, T, P, -1.0);
  for (int e = threadIdx.x; e < (N << lgT); e += NSL_NT) {
    const int n = e >> lgT, t = e & (T - 1), kx = kx0 + t;
    const int cy = compact_of(n, g);
    if (cy >= 0 && kx < g.hx) X4[(outb + cy) * g.hp + kx] = lds[e];
  }
}

// X4[c][z][cy][kx] -> OUT[c][cz][cy][kx] (compact, retained kz only), times inv = 1/N^3 (nslab.js:121-127).
// Grid: (ntile*by, ncomp).
template <int TN>
__global__ __launch_bounds__(NSL_NT) void k_zfwd(const double2 *__restrict__ X4, double2 *__restrict__ OUT,
                                                  const Geo g, const DPlan P, int lgT, double inv,
                                                  const Ctl *__restrict__ ctl) {
  if (ctl != nullptr && ctl->active == 0) return;
  extern __shared__ double2 lds[];
  const int T = 1 << lgT, N = g.N;
  const int ntile = (g.hx + T - 1) >> lgT;
  const int tile = blockIdx.x % ntile, cy = blockIdx.x / ntile, c = blockIdx.y;
  const int kx0 = tile << lgT;
  for (int e = threadIdx.x; e < (N << lgT); e += NSL_NT) {
    const int n = e >> lgT, t = e & (T - 1), kx = kx0 + t;
    lds[e] = kx < g.hx ? X4[(((size_t)c * N + n) * g.by + cy) * g.hp + kx] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  lds_fft<TN>(lds, T, P, -1.0);
  for (int e = threadIdx.x; e < (N << lgT); e += NSL_NT) {
    const int n = e >> lgT, t = e & (T - 1), kx = kx0 + t;
    const int cz = compact_of(n, g);
    if (cz >= 0 && kx < g.hx) {
      const double2 v = lds[e];
      OUT[(size_t)c * g.D + ((size_t)cz * g.by + cy) * g.hp + kx] = make_double2(v.x * inv, v.y * inv);
    }
  }
}

// ------------------------------------------------------------------------------------------------ spectral sweep
// Wavenumbers of compact slot m; false for the padding slots that round each kx row up to the pitch hp.
__device__ __forceinline__ bool mode_k(long long m, const Geo &g, double &kx, double &ky, double &kz) {
  const int ix = (int)(m % g.hp);
  const long long r = m / g.hp;
  const int cy = (int)(r % g.by), cz = (int)(r / g.by);
  kx = (double)ix; ky = kwv(full_of(cy, g), g.N); kz = kwv(full_of(cz, g), g.N);
  return ix < g.hx;
}

#define NSL_MPT 4   // modes per thread in the spectral sweeps

// STAGE 1: ACC = K, T = S + dt/2 K (dt chosen from this rhs's max|u|)   (nslab.js:316)
// STAGE 2: ACC += 2K, T = S + dt/2 K                                     (nslab.js:318)
// STAGE 3: ACC += 2K, T = S + dt K                                       (nslab.js:320)
// STAGE 4: S += dt/6 (ACC + K), and E, Z, P of the new S                  (nslab.js:322-325)
// K = P[NL] - nu k^2 U (nslab.js:235, 240), with the rhs sums of nslab.js:238-244 into partials[block][8].
// The update of one compact mode m, in two halves so a caller can issue the loads of several modes before it
// applies any: upd_load reads the mode's state operands (U; S, which is U itself at stage 1; ACC from stage 2 on),
// upd_apply projects the nonlinear term n0..n2 as the z-forward produced it, forms K, writes the stage's outputs and
// adds the mode's rhs sums (wavenumbers kx, ky, kz, Hermitian weight h) to acc. Shared by k_update and the fused
// z-forward.
struct UpdIn {
  double2 u[3], s[3], a[3];
};
template <int STAGE, class E>
__device__ __forceinline__ UpdIn upd_load(size_t m, long long D, const E *U, const E *S, const E *ACC) {
  UpdIn r;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    r.u[c] = ldc(U[c * D + m]);
    r.s[c] = STAGE == 1 ? r.u[c] : ldc(S[c * D + m]);   // stage 1 evaluates the rhs at S itself
    r.a[c] = STAGE == 1 ? make_double2(0.0, 0.0) : ldc(ACC[c * D + m]);
  }
  return r;
}
template <int STAGE, class E>
__device__ __forceinline__ void upd_apply(const UpdIn &in, size_t m, long long D, double kx, double ky, double kz,
                                          double h, double2 n0, double2 n1, double2 n2, E *S, E *ACC, E *TS,
                                          double nu, double dt, double (&acc)[8]) {
  const double k2 = kx * kx + ky * ky + kz * kz;
  if (k2 == 0) {
    n0 = n1 = n2 = make_double2(0.0, 0.0);
  } else {   // project (nslab.js:194-202)
    const double dr = (kx * n0.x + ky * n1.x + kz * n2.x) / k2, di = (kx * n0.y + ky * n1.y + kz * n2.y) / k2;
    n0.x -= kx * dr; n0.y -= kx * di; n1.x -= ky * dr; n1.y -= ky * di; n2.x -= kz * dr; n2.y -= kz * di;
  }
  const double2 u0 = in.u[0], u1 = in.u[1], u2 = in.u[2];
  double dot = 0, e = 0;
  dot += u0.x * n0.x + u0.y * n0.y; e += u0.x * u0.x + u0.y * u0.y;
  dot += u1.x * n1.x + u1.y * n1.y; e += u1.x * u1.x + u1.y * u1.y;
  dot += u2.x * n2.x + u2.y * n2.y; e += u2.x * u2.x + u2.y * u2.y;
  const double nk = nu * k2;
  const double2 K0 = make_double2(n0.x - nk * u0.x, n0.y - nk * u0.y);
  const double2 K1 = make_double2(n1.x - nk * u1.x, n1.y - nk * u1.y);
  const double2 K2 = make_double2(n2.x - nk * u2.x, n2.y - nk * u2.y);
  acc[0] += h * dot; acc[1] += h * k2 * dot; acc[2] += h * e; acc[3] += h * k2 * e; acc[4] += h * k2 * k2 * e;
  const double2 s0 = in.s[0], s1 = in.s[1], s2 = in.s[2];
  if (STAGE == 1 || STAGE == 2 || STAGE == 3) {
    const double f = STAGE == 3 ? dt : 0.5 * dt;
    stc(TS[m], make_double2(s0.x + f * K0.x, s0.y + f * K0.y));
    stc(TS[D + m], make_double2(s1.x + f * K1.x, s1.y + f * K1.y));
    stc(TS[2 * D + m], make_double2(s2.x + f * K2.x, s2.y + f * K2.y));
    if (STAGE == 1) {
      stc(ACC[m], K0); stc(ACC[D + m], K1); stc(ACC[2 * D + m], K2);
    } else {
      const double2 a0 = in.a[0], a1 = in.a[1], a2 = in.a[2];
      stc(ACC[m], make_double2(a0.x + 2 * K0.x, a0.y + 2 * K0.y));
      stc(ACC[D + m], make_double2(a1.x + 2 * K1.x, a1.y + 2 * K1.y));
      stc(ACC[2 * D + m], make_double2(a2.x + 2 * K2.x, a2.y + 2 * K2.y));
    }
  } else {
    const double f = dt / 6;
    const double2 a0 = in.a[0], a1 = in.a[1], a2 = in.a[2];
    // the new state as stored (rounded to the storage precision), so its E, Z, P match the state itself
    const double2 r0 = rnd<E>(make_double2(s0.x + f * (a0.x + K0.x), s0.y + f * (a0.y + K0.y)));
    const double2 r1 = rnd<E>(make_double2(s1.x + f * (a1.x + K1.x), s1.y + f * (a1.y + K1.y)));
    const double2 r2 = rnd<E>(make_double2(s2.x + f * (a2.x + K2.x), s2.y + f * (a2.y + K2.y)));
    stc(S[m], r0); stc(S[D + m], r1); stc(S[2 * D + m], r2);
    const double en = h * (r0.x * r0.x + r0.y * r0.y + r1.x * r1.x + r1.y * r1.y + r2.x * r2.x + r2.y * r2.y);
    acc[5] += en; acc[6] += k2 * en; acc[7] += k2 * k2 * en;
  }
}

// Zeros where upd_apply<STAGE> writes, for a padding column kx >= hx of the compact rows (its caller writes whole
// 128-byte lines).
template <int STAGE, class E>
__device__ __forceinline__ void upd_zero(size_t m, long long D, E *S, E *ACC, E *TS) {
  const double2 z = make_double2(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (STAGE == 4) {
      stc(S[c * D + m], z);
    } else {
      stc(TS[c * D + m], z);
      stc(ACC[c * D + m], z);
    }
  }
}

template <int STAGE>
// U aliases S (stage 1) or TS (stages 2-4): each mode is read before it is written, by the same thread.
__global__ __launch_bounds__(NSL_NT) void k_update(const double2 *__restrict__ NL, const double2 *U,
                                                    double2 *S, double2 *__restrict__ ACC,
                                                    double2 *TS, const Geo g,
                                                    double *__restrict__ partials, Ctl *__restrict__ ctl, size_t sb) {
  __shared__ double scratch[8 * NSL_NT];
  NL = at_slot(NL, sb); U = at_slot(U, sb); S = at_slot(S, sb); ACC = at_slot(ACC, sb); TS = at_slot(TS, sb);
  partials = at_slot(partials, sb); ctl = at_slot(ctl, sb);
  if (ctl->active == 0) return;
  const double nu = ctl->nu;
  const double dt = STAGE == 1 ? choose_dt(ctl, sqrt(__longlong_as_double((long long)ctl->u2))) : ctl->dt;
  const long long D = g.D;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // Tnl Pspec E Z Pal nE nZ nP
  const long long m0 = (long long)blockIdx.x * NSL_NT * NSL_MPT + threadIdx.x;
#pragma unroll
  for (int q = 0; q < NSL_MPT; q++) {
    const long long m = m0 + (long long)q * NSL_NT;
    if (m >= D) break;
    double kx, ky, kz;
    if (!mode_k(m, g, kx, ky, kz)) {   // padding column: zeros, so the outputs' 128-byte lines are written whole
      upd_zero<STAGE>(m, D, S, ACC, TS);
      continue;
    }
    const double h = (kx == 0 || 2 * kx == g.N) ? 1.0 : 2.0;
    const UpdIn in = upd_load<STAGE>(m, D, U, S, ACC);
    upd_apply<STAGE>(in, m, D, kx, ky, kz, h, NL[m], NL[D + m], NL[2 * D + m], S, ACC, TS, nu, dt, acc);
  }
  block_sum<8>(acc, scratch, partials + (size_t)blockIdx.x * 8);
}

// Sums the per-block partials in a fixed order and books the rhs diagnostics. mode 0..3: stage mode+1 of the step
// in flight; mode 4: `last` (rhs of the current state) and, when rec != null, the end of the step: budgets
// (nslab.js:326-329), t += dt, step++, one StepRecDev.
// Run by one workgroup of NT threads (any NT), which plays NSL_NT virtual threads: virtual thread v adds rows v,
// v + NSL_NT, ... and the rows' totals follow block_sum<8, NSL_NT>'s tree, CH components at a time, so every
// caller gets the same bits whatever its own workgroup size. scratch: CH * NSL_NT doubles of LDS.
template <int NT, int CH = 4>
__device__ __forceinline__ void fin_rhs(const double *__restrict__ partials, int nblk, int mode,
                                        StepRecDev *__restrict__ rec, Ctl *__restrict__ ctl, double *scratch) {
  static_assert(CH == 1 || CH == 2 || CH == 4 || CH == 8, "components per round divide 8");
  constexpr int V = NSL_NT, KV = (V + NT - 1) / NT, P2 = pow2_ceil(V);
  double acc[KV][8];
#pragma unroll
  for (int k = 0; k < KV; k++) {
    const int v = threadIdx.x + k * NT;
#pragma unroll
    for (int i = 0; i < 8; i++) acc[k][i] = 0.0;
    if (v < V)
      for (int b = v; b < nblk; b += V)
#pragma unroll
        for (int i = 0; i < 8; i++) acc[k][i] += partials[(size_t)b * 8 + i];
  }
  double tot[8];
#pragma unroll
  for (int h = 0; h < 8 / CH; h++) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KV; k++) {
      const int v = threadIdx.x + k * NT;
      if (v < V)
#pragma unroll
        for (int i = 0; i < CH; i++) scratch[i * V + v] = acc[k][CH * h + i];
    }
    __syncthreads();
    for (int w = P2 / 2; w > 0; w >>= 1) {
#pragma unroll
      for (int k = 0; k < KV; k++) {
        const int v = threadIdx.x + k * NT;
        if (v < w && v + w < V)
#pragma unroll
          for (int i = 0; i < CH; i++) scratch[i * V + v] += scratch[i * V + v + w];
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < CH; i++) tot[CH * h + i] = scratch[i * V];
  }
  if (threadIdx.x != 0) return;
  ctl->pend = 0;
  const double nu = ctl->nu;
  double *d = ctl->diag[mode];
  const double Z = tot[3];
  d[0] = 0.5 * tot[2]; d[1] = 0.5 * Z; d[2] = 0.5 * tot[4]; d[3] = nu * Z; d[4] = tot[0]; d[5] = tot[1];
  d[6] = sqrt(__longlong_as_double((long long)ctl->om2));
  d[7] = sqrt(__longlong_as_double((long long)ctl->u2));
  ctl->om2 = 0; ctl->u2 = 0;
  if (mode == 3) { ctl->nrm[0] = 0.5 * tot[5]; ctl->nrm[1] = 0.5 * tot[6]; ctl->nrm[2] = 0.5 * tot[7]; }
  if (mode == 4 && rec != nullptr) {
    const double dt = ctl->dt;
    double eps[4], zr[4];
    for (int s = 0; s < 4; s++) { eps[s] = ctl->diag[s][3]; zr[s] = ctl->diag[s][5] - 2 * nu * ctl->diag[s][2]; }
    const double E0 = ctl->diag[0][0], Z0 = ctl->diag[0][1], nE = ctl->nrm[0], nZ = ctl->nrm[1];
    ctl->t += dt; ctl->step++;
    const double dEpred = -dt / 6 * (eps[0] + 2 * eps[1] + 2 * eps[2] + eps[3]);
    const double dZpred = dt / 6 * (zr[0] + 2 * zr[1] + 2 * zr[2] + zr[3]);
    const double ebal = fabs((nE - E0) - dEpred) / fmax(ctl->E0ref, 1e-300);
    const double zbal = fabs((nZ - Z0) - dZpred) / fmax(fmax(nZ, Z0), 1e-300);
    ctl->ebal = ebal; ctl->zbal = zbal;
    StepRecDev &R = rec[ctl->nChunk];
    R.t = ctl->t; R.dt = dt; R.step = ctl->step; R.ebal = ebal; R.zbal = zbal;
    for (int i = 0; i < 8; i++) { R.last[i] = d[i]; R.d0[i] = ctl->diag[0][i]; }
    ctl->nChunk++;
    ctl->nDone++;
  }
}

// The rhs's bookkeeping as a kernel of its own. need_pend = 0: right after an update that does not defer it (the
// generic and two-kernel paths); need_pend = 1: whatever a fused update left pending (the end of a chunk of steps,
// a rhs outside a step), nothing otherwise.
__global__ __launch_bounds__(NSL_NT) void k_finalize(const double *__restrict__ partials, int nblk, int mode,
                                                      StepRecDev *__restrict__ rec, Ctl *__restrict__ ctl, size_t sb,
                                                      int need_pend) {
  __shared__ double scratch[4 * NSL_NT];
  partials = at_slot(partials, sb); rec = at_slot(rec, sb); ctl = at_slot(ctl, sb);
  if (ctl->active == 0 || (need_pend && ctl->pend == 0)) return;
  fin_rhs<NSL_NT>(partials, nblk, mode, rec, ctl, scratch);
}

// Start of a step: first the end of the previous step, if a fused update left it pending (nfin partials, mode 4,
// one StepRecDev), then the stop predicate of run() (nslab.js:334) and dt of this step (nslab.js:314).
__global__ __launch_bounds__(NSL_NT) void k_begin_step(Ctl *ctl, const double *__restrict__ partials, int nfin,
                                                       StepRecDev *__restrict__ rec, size_t sb) {
  __shared__ double scratch[4 * NSL_NT];
  ctl = at_slot(ctl, sb); partials = at_slot(partials, sb); rec = at_slot(rec, sb);
  if (nfin > 0 && ctl->pend != 0) fin_rhs<NSL_NT>(partials, nfin, 4, rec, ctl, scratch);
  if (threadIdx.x != 0) return;
  const bool act = ctl->nDone < ctl->nTarget && !(ctl->hasTEnd && ctl->t >= ctl->tEnd - 1e-12);
  ctl->active = act ? 1 : 0;
  if (act) {
    ctl->dt = choose_dt(ctl, ctl->diag[4][7]);
    for (int i = 0; i < 8; i++) ctl->diag[0][i] = ctl->diag[4][i];
  }
}
__global__ void k_set_active(Ctl *ctl, int v, size_t sb) { at_slot(ctl, sb)->active = v; }
__global__ void k_chunk_begin(Ctl *ctl, size_t sb) { at_slot(ctl, sb)->nChunk = 0; }

// ------------------------------------------------------------------------------------------------ set-up kernels
// Physical initial fields (nslab.js:251-266): kind 0 tgv, 1 tgv2d, 2 abc into phys[0..2] (u, v, w);
// kind 3 tubes: omega_x into phys[0], phys[1..2] = 0; kind 4 (tube pair along z): omega into phys[0..2].
__global__ void k_ic_phys(double *__restrict__ phys, int N, int kind, double p0, double p1, double p2, double p3) {
  const size_t N3 = (size_t)N * N * N;
  const double h = NSL_TWO_PI / N, PI = 3.141592653589793;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < N3; n += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(n % N), j = (int)((n / N) % N), l = (int)(n / ((size_t)N * N));
    const double x = i * h, y = j * h, z = l * h;
    double u = 0, v = 0, w = 0;
    if (kind == 0) { u = sin(x) * cos(y) * cos(z); v = -cos(x) * sin(y) * cos(z); w = 0; }
    else if (kind == 1) { u = sin(x) * cos(y); v = -cos(x) * sin(y); w = 0; }
    else if (kind == 2) { u = p0 * sin(z) + p2 * cos(y); v = p1 * sin(x) + p0 * cos(z); w = p2 * sin(y) + p1 * cos(x); }
    else if (kind == 4) {
      // two Gaussian-core tubes along z, circulation +-p0, core radius p1, centres (pi +- (p2 + 0.1 cos z), pi);
      // omega along the centreline tangent (dx_c/dz, 0, 1), nearest periodic image over a 3x3 lattice
      const double G = p0, sig = p1, d0 = p2, a = 0.1;
      double wx = 0, wz = 0;
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        const double xc = PI + sgn * (d0 + a * cos(z)), yc = PI;
        double best = INFINITY;
        for (int px = -1; px <= 1; px++)
          for (int py = -1; py <= 1; py++) {
            const double dx = x - xc + px * NSL_TWO_PI, dy = y - yc + py * NSL_TWO_PI;
            best = fmin(best, dx * dx + dy * dy);
          }
        const double om = sgn * G / (PI * sig * sig) * exp(-best / (sig * sig));
        wz += om;
        wx += -sgn * a * sin(z) * om;
      }
      u = wx; v = 0; w = wz;
    } else {
      const double amp = p0, sig = p1, d0 = p2, del = p3;
      const double d = d0 + del * cos(x);
      double s = 0;
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        const double yc = PI + sgn * d, zc = PI;
        double best = INFINITY;
        for (int py = -1; py <= 1; py++)
          for (int pz = -1; pz <= 1; pz++) {
            const double dy = y - yc + py * NSL_TWO_PI, dz = z - zc + pz * NSL_TWO_PI;
            best = fmin(best, dy * dy + dz * dz);
          }
        s += -sgn * amp * exp(-best / (sig * sig));
      }
      u = s; v = 0; w = 0;
    }
    phys[n] = u; phys[N3 + n] = v; phys[2 * N3 + n] = w;
  }
}

__device__ __forceinline__ void project_mode(double kx, double ky, double kz, double k2, double2 &a, double2 &b,
                                             double2 &c) {
  if (k2 == 0) { a = b = c = make_double2(0.0, 0.0); return; }
  const double dr = (kx * a.x + ky * b.x + kz * c.x) / k2, di = (kx * a.y + ky * b.y + kz * c.y) / k2;
  a.x -= kx * dr; a.y -= kx * di; b.x -= ky * dr; b.y -= ky * di; c.x -= kz * dr; c.y -= kz * di;
}

// project(U, true) over compact storage (the mask is implicit: only retained modes are stored).
__global__ void k_project(double2 *__restrict__ U, const Geo g) {
  const long long D = g.D;
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < D; m += (long long)gridDim.x * blockDim.x) {
    double kx, ky, kz;
    if (!mode_k(m, g, kx, ky, kz)) continue;
    const double k2 = kx * kx + ky * ky + kz * kz;
    double2 a = U[m], b = U[D + m], c = U[2 * D + m];
    project_mode(kx, ky, kz, k2, a, b, c);
    U[m] = a; U[D + m] = b; U[2 * D + m] = c;
  }
}

// Random solenoidal field, the draws (nslab.js:273-283): mode (kz, ky, kx) of the mask, in the reference's loop order
// over the masked modes (the compact storage order: cz, then cy, then kx), takes 12 Mulberry32 numbers (3 components
// x 2 parts x 2 draws, Box-Muller), amplitude sqrt(k^4 exp(-2 (k/k0)^2) / (4 pi k^2)); the modes that do not draw
// (k = 0, and on kx = 0 the half kz < 0 or kz = 0, ky < 0) are left to the conjugate fill of their partner. Mulberry32's
// state after n draws is a0 + n * 0x6D2B79F5 (mod 2^32), so each mode starts from its own draw count: with hx kx per
// row and the kx = 0 mode drawing only for kz > 0 or kz = 0 < ky, a row r = cz * by + cy starts after
// r (hx - 1) + (rows before r whose kx = 0 mode draws) draws of 12 numbers. dst (3 components of D slots) is zero
// on entry. One thread per mode (cz, cy, kx < hx).
// Mulberry32 (nslab.js:161): a = a + 0x6D2B79F5, then a 32-bit hash of a.
__device__ __forceinline__ double mb32_next_dev(unsigned &a) {
  a += 0x6D2B79F5u;
  unsigned t = (a ^ (a >> 15)) * (1u | a);
  t = (t + (t ^ (t >> 7)) * (61u | t)) ^ t;
  return (double)(t ^ (t >> 14)) / 4294967296.0;
}
__global__ void k_ic_random(double2 *__restrict__ dst, const Geo g, unsigned a0, double k0) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)g.by * g.by * g.hx) return;
  const int i = (int)(e % g.hx);
  const long long r = e / g.hx;
  const int cy = (int)(r % g.by), cz = (int)(r / g.by), N = g.N, kc = g.hy - 1;
  const int l = full_of(cz, g), j = full_of(cy, g);
  const double kx = i, ky = kwv(j, N), kz = kwv(l, N);
  const double k2 = kx * kx + ky * ky + kz * kz;
  const bool draw0 = kz > 0 || (kz == 0 && ky > 0);   // this row's kx = 0 mode draws
  if (k2 == 0 || (i == 0 && !draw0)) return;
  // rows before r whose kx = 0 mode draws: cz = 0 rows with 0 < ky (cy 1..kc), every row of 0 < kz (cz 1..kc)
  const long long d0 = cz == 0 ? (long long)max(0, min(cy, kc + 1) - 1)
                                : cz <= kc ? (long long)kc + (long long)(cz - 1) * g.by + cy : (long long)kc + (long long)kc * g.by;
  const long long before = r * (g.hx - 1) + d0 + (i == 0 ? 0 : (draw0 ? 1 : 0) + (i - 1));
  unsigned a = a0 + (unsigned)((unsigned long long)before * 12ull * 0x6D2B79F5ull);
  const double k = sqrt(k2), amp = sqrt(pow(k, 4.0) * exp(-2 * (k / k0) * (k / k0)) / (4 * M_PI * k2));
  const size_t m = ((size_t)cz * g.by + cy) * g.hp + i;
  double2 v[3];
  for (int c = 0; c < 3; c++) {
    double p[2];
    for (int part = 0; part < 2; part++) {
      double ra = mb32_next_dev(a);
      ra = fmax(ra, 1e-12);
      const double rb = mb32_next_dev(a);
      p[part] = amp * (sqrt(-2 * log(ra)) * cos(NSL_TWO_PI * rb));
    }
    v[c] = make_double2(p[0], p[1]);
    dst[(size_t)c * g.D + m] = v[c];
  }
  if (i == 0) {   // the conjugate partner (a mode that does not draw: no other thread writes it)
    const int czm = compact_of((N - l) % N, g), cym = compact_of((N - j) % N, g);
    const size_t mm = ((size_t)czm * g.by + cym) * g.hp;
    for (int c = 0; c < 3; c++) dst[(size_t)c * g.D + mm] = make_double2(v[c].x, -v[c].y);
  }
}

// Tubes (nslab.js:267-272): W = project(omega-hat); U = (i k x W) / k^2 (k = 0 -> /1); S = project(U).
__global__ void k_tubes_finish(const double2 *__restrict__ W, double2 *__restrict__ S, const Geo g) {
  const long long D = g.D;
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < D; m += (long long)gridDim.x * blockDim.x) {
    double kx, ky, kz;
    if (!mode_k(m, g, kx, ky, kz)) continue;
    double k2 = kx * kx + ky * ky + kz * kz;
    double2 a = W[m], b = W[D + m], c = W[2 * D + m];
    project_mode(kx, ky, kz, k2, a, b, c);
    double pr = ky * c.x - kz * b.x, pi = ky * c.y - kz * b.y;
    double2 u0 = make_double2(-pi, pr);
    pr = kz * a.x - kx * c.x; pi = kz * a.y - kx * c.y;
    double2 u1 = make_double2(-pi, pr);
    pr = kx * b.x - ky * a.x; pi = kx * b.y - ky * a.y;
    double2 u2 = make_double2(-pi, pr);
    const double kd = k2 == 0 ? 1.0 : k2;
    u0.x /= kd; u0.y /= kd; u1.x /= kd; u1.y /= kd; u2.x /= kd; u2.y /= kd;
    project_mode(kx, ky, kz, k2, u0, u1, u2);
    S[m] = u0; S[D + m] = u1; S[2 * D + m] = u2;
  }
}

// S += T over n complex values (a perturbation added to an initial field).
__global__ void k_axpy(double2 *__restrict__ S, const double2 *__restrict__ T, long long n) {
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (long long)gridDim.x * blockDim.x) {
    const double2 a = S[m], b = T[m];
    S[m] = make_double2(a.x + b.x, a.y + b.y);
  }
}

// norms(U) partial sums (nslab.js:214-222) into partials[block][8] = {-,-,E,Z,P,-,-,-} so k_finalize-style sums
// apply; the totals are reduced by k_norms_finalize.
__global__ __launch_bounds__(NSL_NT) void k_norms(const double2 *__restrict__ U, const Geo g,
                                                   double *__restrict__ partials) {
  __shared__ double scratch[3 * NSL_NT];
  const long long D = g.D;
  double acc[3] = {0, 0, 0};
  const long long m0 = (long long)blockIdx.x * NSL_NT * NSL_MPT + threadIdx.x;
#pragma unroll
  for (int q = 0; q < NSL_MPT; q++) {
    const long long m = m0 + (long long)q * NSL_NT;
    if (m >= D) break;
    double kx, ky, kz;
    if (!mode_k(m, g, kx, ky, kz)) continue;
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double2 a = U[m], b = U[D + m], c = U[2 * D + m];
    const double h = (kx == 0 || 2 * kx == g.N) ? 1.0 : 2.0;
    const double e = h * (a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y + c.x * c.x + c.y * c.y);
    acc[0] += e; acc[1] += k2 * e; acc[2] += k2 * k2 * e;
  }
  block_sum<3>(acc, scratch, partials + (size_t)blockIdx.x * 3);
}

__global__ __launch_bounds__(NSL_NT) void k_norms_finalize(const double *__restrict__ partials, int nblk,
                                                            double *__restrict__ out3) {
  __shared__ double scratch[3 * NSL_NT];
  __shared__ double tot[3];
  double acc[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += NSL_NT)
    for (int i = 0; i < 3; i++) acc[i] += partials[(size_t)b * 3 + i];
  block_sum<3>(acc, scratch, tot);
  if (threadIdx.x == 0) { out3[0] = 0.5 * tot[0]; out3[1] = 0.5 * tot[1]; out3[2] = 0.5 * tot[2]; }
}

// S *= sqrt(E0 / E) with E = out3[0] (random IC normalisation, nslab.js:285).
__global__ void k_scale_energy(double2 *__restrict__ U, long long n, const double *__restrict__ nrm, double E0) {
  const double sc = sqrt(E0 / nrm[0]);
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (long long)gridDim.x * blockDim.x) {
    const double2 v = U[m];
    U[m] = make_double2(v.x * sc, v.y * sc);
  }
}

// Batched 1D lines (fft1 export): data[line][n] split re/im, T lines per workgroup.
template <int TN>
__global__ __launch_bounds__(NSL_NT) void k_fft_lines(double *__restrict__ re, double *__restrict__ im,
                                                       long long lines, const DPlan P, int T, double sign) {
  extern __shared__ double2 lds[];
  const int N = P.N;
  const long long l0 = (long long)blockIdx.x * T;
  for (int e = threadIdx.x; e < N * T; e += NSL_NT) {
    const int t = e / N, n = e - t * N;
    const long long l = l0 + t;
    lds[n * T + t] = l < lines ? make_double2(re[l * N + n], im[l * N + n]) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  lds_fft<TN>(lds, T, P, sign);
  for (int e = threadIdx.x; e < N * T; e += NSL_NT) {
    const int t = e / N, n = e - t * N;
    const long long l = l0 + t;
    if (l < lines) { const double2 v = lds[n * T + t]; re[l * N + n] = v.x; im[l * N + n] = v.y; }
  }
}

}  // namespace nsl
