// Device-side building blocks of the NSLab gfx950 solver: geometry of the 2/3-truncated spectral storage, the
// LDS-staged mixed-radix Stockham FFT engine, and block reductions.
//
// Arithmetic is IEEE binary64 throughout. The library is compiled with -ffp-contract=off so every product and sum
// rounds as the reference's JavaScript does (pocket-wind-tunnel/src/nslab.js); the butterflies below use the same
// radix-2/3/4 decimation-in-frequency Stockham formulas and factor order as nslab.js:46-92, so the transforms differ
// from the reference only in the last-ulp rounding of the twiddle table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#define NSL_NT 256          // threads per workgroup (4 wave64)
#define NSL_MAXST 16        // max FFT stages (radix >= 2 -> N <= 2^16)

namespace nsl {

// Retained-mode geometry. Spectral data along kx keeps hx modes 0..hx-1; along ky and kz it keeps `by` modes stored
// compactly: compact index c < hy is full index c (k = c), c >= hy is full index c + N - by (k = c - by).
// The solver state uses hx = kc+1, by = 2kc+1, hy = kc+1 (kc = floor(N/3), the 2/3 rule of nslab.js:176);
// an untruncated transform uses hx = N/2+1, by = N, hy = N/2.
struct Geo {
  int N, NH, hx, by, hy;
  int hp;        // row pitch along kx (hx rounded up to 8 complex = 128 bytes for the solver state and pencils)
  long long D;   // slots per component in compact storage = by*by*hp
};

struct DPlan {
  int N, nst;
  int r[NSL_MAXST];       // radix of each stage (4s, then a 2, then 3s: nslab.js:28-35)
  int tw[NSL_MAXST];      // offset of each stage's twiddles in the table
  const double2 *twd;     // (cos, sin)(2*pi*j*p/n) at [(j-1)*m + p] per stage
};

__device__ __forceinline__ int full_of(int c, const Geo &g) { return c < g.hy ? c : c + g.N - g.by; }
__device__ __forceinline__ int compact_of(int f, const Geo &g) {
  return f < g.hy ? f : (f >= g.N - g.by + g.hy ? f - (g.N - g.by) : -1);
}
// The dispatcher deals workgroups to the 8 XCDs round-robin, and each XCD has its own L2. This maps block ids so
// each XCD works on one contiguous range of logical tiles: tiles that share inputs then share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  const int per = nblocks / 8, rem = nblocks % 8, x = bid % 8, l = bid / 8;
  return x * per + min(x, rem) + l;
}

// signed wavenumber of full index f (nslab.js:180: j <= N/2 ? j : j - N)
__device__ __forceinline__ double kwv(int f, int N) { return (double)(f <= (N >> 1) ? f : f - N); }

// ------------------------------------------------------------------------------------------------ butterflies
template <int R>
__device__ __forceinline__ void bfly(double2 *v, const double2 *__restrict__ tw, int m, double sign);

template <>
__device__ __forceinline__ void bfly<2>(double2 *v, const double2 *__restrict__ tw, int m, double sign) {
  const double2 a = v[0], b = v[1];
  v[0] = make_double2(a.x + b.x, a.y + b.y);
  const double tr = a.x - b.x, ti = a.y - b.y;
  const double2 w = tw[0];
  const double c = w.x, d = sign * w.y;
  v[1] = make_double2(tr * c - ti * d, tr * d + ti * c);
}

template <>
__device__ __forceinline__ void bfly<4>(double2 *v, const double2 *__restrict__ tw, int m, double sign) {
  const double2 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
  const double t0r = a0.x + a2.x, t0i = a0.y + a2.y, t1r = a0.x - a2.x, t1i = a0.y - a2.y;
  const double t2r = a1.x + a3.x, t2i = a1.y + a3.y, t3r = a1.x - a3.x, t3i = a1.y - a3.y;
  const double2 w1 = tw[0], w2 = tw[m], w3 = tw[2 * m];
  const double c1 = w1.x, d1 = sign * w1.y, c2 = w2.x, d2 = sign * w2.y, c3 = w3.x, d3 = sign * w3.y;
  v[0] = make_double2(t0r + t2r, t0i + t2i);
  double br = t1r - sign * t3i, bi = t1i + sign * t3r;          // t1 + sign*i*t3
  v[1] = make_double2(br * c1 - bi * d1, br * d1 + bi * c1);
  br = t0r - t2r; bi = t0i - t2i;
  v[2] = make_double2(br * c2 - bi * d2, br * d2 + bi * c2);
  br = t1r + sign * t3i; bi = t1i - sign * t3r;                 // t1 - sign*i*t3
  v[3] = make_double2(br * c3 - bi * d3, br * d3 + bi * c3);
}

template <>
__device__ __forceinline__ void bfly<3>(double2 *v, const double2 *__restrict__ tw, int m, double sign) {
  const double S3 = -sign * 0.8660254037844386;
  const double2 a0 = v[0], a1 = v[1], a2 = v[2];
  const double sr = a1.x + a2.x, si = a1.y + a2.y, ur = a0.x - 0.5 * sr, ui = a0.y - 0.5 * si;
  const double vr = S3 * (a1.x - a2.x), vi = S3 * (a1.y - a2.y);
  const double2 w1 = tw[0], w2 = tw[m];
  const double c1 = w1.x, d1 = sign * w1.y, c2 = w2.x, d2 = sign * w2.y;
  v[0] = make_double2(a0.x + sr, a0.y + si);
  double br = ur + vi, bi = ui - vr;                            // u - i*v
  v[1] = make_double2(br * c1 - bi * d1, br * d1 + bi * c1);
  br = ur - vi; bi = ui + vr;                                   // u + i*v
  v[2] = make_double2(br * c2 - bi * d2, br * d2 + bi * c2);
}

// One Stockham stage over T interleaved lines held in LDS as buf[n*T + t]. Every thread first loads the inputs of
// its butterflies into registers, the workgroup synchronises, then writes the outputs to their autosorted slots:
// the transform runs in place in one LDS buffer. TN bounds T*N (register footprint per thread).
template <int R, int TN>
__device__ __forceinline__ void fft_stage(double2 *__restrict__ buf, int T, int N, int s,
                                          const double2 *__restrict__ tw, double sign) {
  const int NB = N / R, m = NB / s, total = NB * T;
  constexpr int K = (TN / R + NSL_NT - 1) / NSL_NT;
  double2 v[K][R];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int idx = threadIdx.x + k * NSL_NT;
    if (idx < total) {
      const int bb = idx / T, t = idx - bb * T;
#pragma unroll
      for (int j = 0; j < R; j++) v[k][j] = buf[(bb + j * NB) * T + t];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int idx = threadIdx.x + k * NSL_NT;
    if (idx < total) {
      const int bb = idx / T, t = idx - bb * T;
      const int p = bb / s, q = bb - p * s;
      bfly<R>(v[k], tw + p, m, sign);
      const int o0 = q + s * R * p;
#pragma unroll
      for (int j = 0; j < R; j++) buf[(o0 + j * s) * T + t] = v[k][j];
    }
  }
  __syncthreads();
}

// In-place FFT of T lines of length P.N in LDS (layout buf[n*T + t]); sign -1 forward, +1 inverse, unnormalised.
// Caller must __syncthreads() before (data written) — the engine ends synchronised.
template <int TN>
__device__ void lds_fft(double2 *__restrict__ buf, int T, const DPlan &P, double sign) {
  int s = 1;
  for (int g = 0; g < P.nst; g++) {
    const int r = P.r[g];
    const double2 *tw = P.twd + P.tw[g];
    if (r == 4) fft_stage<4, TN>(buf, T, P.N, s, tw, sign);
    else if (r == 2) fft_stage<2, TN>(buf, T, P.N, s, tw, sign);
    else fft_stage<3, TN>(buf, T, P.N, s, tw, sign);
    s *= r;
  }
}

// ------------------------------------------------------------------------------------------------ slots
// A pool holds several solves of one N, slot z's buffers sb bytes after slot z-1's. A pass launched with
// gridDim.z = slots runs every slot at once, each workgroup on the buffers of slot blockIdx.z; sb = 0 with one slot
// is the single-solve case.
// (byte arithmetic on the pointer itself, not through an integer, so the compiler still sees a global-memory
// pointer and emits global rather than flat loads and stores)
template <class T>
__device__ __forceinline__ T *at_slot(T *p, size_t sb) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return p == nullptr ? p : reinterpret_cast<T *>(reinterpret_cast<B *>(p) + (size_t)blockIdx.z * sb);
}

// ------------------------------------------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Max of non-negative doubles over the workgroup into *dst (atomicMax on the IEEE bit pattern: for x >= 0 the
// unsigned order of the bits is the numeric order). Exact, so independent of scheduling.
template <int NT = NSL_NT>
__device__ __forceinline__ void block_atomic_max2(double a, double b, unsigned long long *da,
                                                  unsigned long long *db, double *scratch /* 2*NT/64 */) {
  a = wave_max(a); b = wave_max(b);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) { scratch[2 * wid] = a; scratch[2 * wid + 1] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ma = scratch[0], mb = scratch[1];
    for (int w = 1; w < NT / 64; w++) { ma = fmax(ma, scratch[2 * w]); mb = fmax(mb, scratch[2 * w + 1]); }
    atomicMax(da, (unsigned long long)__double_as_longlong(ma));
    atomicMax(db, (unsigned long long)__double_as_longlong(mb));
  }
}

// Max of two non-negative doubles over the workgroup, stored (not atomically combined) at dst[0], dst[1]: every
// workgroup owns its own slot, so no two workgroups contend for one address.
__device__ __forceinline__ void block_max2_store(double a, double b, double *dst, double *scratch) {
  a = wave_max(a); b = wave_max(b);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) { scratch[2 * wid] = a; scratch[2 * wid + 1] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ma = scratch[0], mb = scratch[1];
    for (int w = 1; w < NSL_NT / 64; w++) { ma = fmax(ma, scratch[2 * w]); mb = fmax(mb, scratch[2 * w + 1]); }
    dst[0] = ma; dst[1] = mb;
  }
}

// Fixed-order tree sum of NV values per thread across the workgroup; thread 0 gets the totals in out[].
// The order depends only on the launch shape, so results are reproducible run to run and batch to batch.
// A workgroup that is not a power of two (48, 96, 192 threads) sums as if padded with zeros to the next one.
__host__ __device__ constexpr int pow2_ceil(int n) { return n <= 1 ? 1 : 2 * pow2_ceil((n + 1) / 2); }
template <int NV, int NT = NSL_NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *scratch /* NT*NV */, double *out) {
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; i++) scratch[i * NT + threadIdx.x] = v[i];
  __syncthreads();
  constexpr int P2 = pow2_ceil(NT);   // a constant: evaluated by the compiler, not called at run time
  for (int w = P2 / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && (int)threadIdx.x + w < NT) {
#pragma unroll
      for (int i = 0; i < NV; i++) scratch[i * NT + threadIdx.x] += scratch[i * NT + threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) out[i] = scratch[i * NT];
  }
}

}  // namespace nsl
