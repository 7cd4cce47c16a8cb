// Size-specialised FFT engine for the pencil passes. The line length N, the lines per workgroup T and the radix
// schedule are compile-time constants, so every index is a shift or a multiply by a constant and every loop unrolls.
// Each stage is a radix-R Stockham decimation-in-frequency step (the same recurrence as nslab.js:46-92, with larger
// radices: an R-point DFT in registers, then the twiddle W_N^(j*p*s) on leg j). The first stage loads its inputs
// straight from global memory and the last stage stores straight to global memory, so LDS is touched only between
// stages: once for N = 16*16, twice for 8*8*8.
#pragma once
#include <type_traits>

#include "nsl_device.hpp"

namespace nsl {
namespace fast {

// The transform arithmetic may fuse a product into the following add (one rounding instead of two); everything
// outside the FFT engines keeps the reference's separate roundings (-ffp-contract=off for the library).
#ifndef NSL_FFT_CONTRACT
#define NSL_FFT_CONTRACT _Pragma("clang fp contract(fast)")
#endif

// The engine is generic in its complex type C: double2 (FP64 solves) or float2 (FP32 solves, which transform in
// single precision as the reference's FP32 runner does). Sc<C> is the scalar type, mk<C> builds a value; every
// expression below is the same for both types, so the double2 instantiation is the FP64 engine (the compiler may
// pick the other product of an a*b - c*d to fuse, a last-bit difference within the 1e-13 parity bound).
template <class C> using Sc = decltype(C{}.x);
template <class C> __device__ __forceinline__ C mk(Sc<C> x, Sc<C> y) { C v; v.x = x; v.y = y; return v; }

template <int SIGN, class C>
__device__ __forceinline__ C mul_tw(C a, C w) {   // a * (w.x + i SIGN w.y)
  NSL_FFT_CONTRACT
  const Sc<C> c = w.x, d = SIGN * w.y;
  return mk<C>(a.x * c - a.y * d, a.x * d + a.y * c);
}
template <class C>
__device__ __forceinline__ C cmul(C a, C b) {
  NSL_FFT_CONTRACT
  return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Twiddles w^j, j = 1..R-1, of one butterfly, w = (cos, sin)(2 pi base / N) and w^j = tw[j * base]. w^1..w^3 and
// w^(4a) are read from the table; every other w^(4a+b) is one product w^(4a) w^b of two table values (about one
// rounding more than a table read): 4 table reads instead of 7 per radix-8 butterfly, 6 instead of 15 per radix-16
// (the table reads, not the products, load the vector memory pipe).
#ifndef NSL_TW_PRODUCTS
#define NSL_TW_PRODUCTS 1
#endif
template <int R, class C>
__device__ __forceinline__ void twiddles(const C *__restrict__ tw, int base, C (&w)[R]) {
#pragma unroll
  for (int j = 1; j < R; j++) {
    if (!NSL_TW_PRODUCTS || j < 4 || (j & 3) == 0) w[j] = tw[j * base];
    else w[j] = cmul(w[j & ~3], w[j & 3]);
  }
}
template <class C> __device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <class C> __device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
// multiply by SIGN * i
template <int SIGN, class C>
__device__ __forceinline__ C mul_si(C a) { return SIGN > 0 ? mk<C>(-a.y, a.x) : mk<C>(a.y, -a.x); }

// In-register DFTs: y_j = sum_k x_k w^(jk), w = exp(SIGN 2 pi i / R), natural order in and out.
template <int R, int SIGN> struct Dft;

template <int SIGN> struct Dft<2, SIGN> {
  template <class C> static __device__ __forceinline__ void run(C *v) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  }
};

template <int SIGN> struct Dft<3, SIGN> {
  template <class C> static __device__ __forceinline__ void run(C *v) {
    NSL_FFT_CONTRACT
    using F = Sc<C>;
    const F S3 = -SIGN * (F)0.8660254037844386, half = (F)0.5;
    const C a0 = v[0], a1 = v[1], a2 = v[2];
    const F sr = a1.x + a2.x, si = a1.y + a2.y, ur = a0.x - half * sr, ui = a0.y - half * si;
    const F vr = S3 * (a1.x - a2.x), vi = S3 * (a1.y - a2.y);
    v[0] = mk<C>(a0.x + sr, a0.y + si);
    v[1] = mk<C>(ur + vi, ui - vr);
    v[2] = mk<C>(ur - vi, ui + vr);
  }
};

template <int SIGN> struct Dft<4, SIGN> {
  template <class C> static __device__ __forceinline__ void run(C *v) {
    const C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    const C st3 = mul_si<SIGN>(t3);
    v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
    v[1] = cadd(t1, st3); v[3] = csub(t1, st3);
  }
};

template <int SIGN> struct Dft<6, SIGN> {
  template <class C> static __device__ __forceinline__ void run(C *v) {
    // n = 2 n1 + n2, j = j1 + 3 j2: DFT3 over n1, twiddle w6^(j1 n2), DFT2 over n2
    using F = Sc<C>;
    C e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
    Dft<3, SIGN>::run(e);
    Dft<3, SIGN>::run(o);
    o[1] = mul_tw<SIGN>(o[1], mk<C>((F)0.5, (F)0.8660254037844386));    // w6
    o[2] = mul_tw<SIGN>(o[2], mk<C>((F)-0.5, (F)0.8660254037844386));   // w6^2
#pragma unroll
    for (int j1 = 0; j1 < 3; j1++) { v[j1] = cadd(e[j1], o[j1]); v[j1 + 3] = csub(e[j1], o[j1]); }
  }
};

template <int SIGN> struct Dft<8, SIGN> {
  template <class C> static __device__ __forceinline__ void run(C *v) {
    NSL_FFT_CONTRACT
    using F = Sc<C>;
    const F H = (F)0.7071067811865476;
    C a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; k++) { a[k] = cadd(v[k], v[k + 4]); b[k] = csub(v[k], v[k + 4]); }
    // b_k *= w8^k, w8 = exp(SIGN i pi/4)
    b[1] = mk<C>(H * (b[1].x - SIGN * b[1].y), H * (b[1].y + SIGN * b[1].x));
    b[2] = mul_si<SIGN>(b[2]);
    b[3] = mk<C>(H * (-b[3].x - SIGN * b[3].y), H * (-b[3].y + SIGN * b[3].x));
    Dft<4, SIGN>::run(a);
    Dft<4, SIGN>::run(b);
#pragma unroll
    for (int m = 0; m < 4; m++) { v[2 * m] = a[m]; v[2 * m + 1] = b[m]; }
  }
};

template <int SIGN> struct Dft<16, SIGN> {
  template <class C> static __device__ __forceinline__ void run(C *v) {
    // n = 4 n1 + n2, j = j1 + 4 j2: DFT4 over n1, twiddle w16^(j1 n2), DFT4 over n2
    using F = Sc<C>;
    const F C1 = (F)0.9238795325112867, S1 = (F)0.3826834323650898, H = (F)0.7071067811865476;
    C A[4][4];   // A[n2][j1]
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
      C c[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Dft<4, SIGN>::run(c);
#pragma unroll
      for (int j1 = 0; j1 < 4; j1++) A[n2][j1] = c[j1];
    }
    // twiddles w16^(j1 n2) = exp(SIGN i 2 pi j1 n2 / 16) for j1, n2 in 1..3
    A[1][1] = mul_tw<SIGN>(A[1][1], mk<C>(C1, S1));
    A[1][2] = mul_tw<SIGN>(A[1][2], mk<C>(H, H));
    A[1][3] = mul_tw<SIGN>(A[1][3], mk<C>(S1, C1));
    A[2][1] = mul_tw<SIGN>(A[2][1], mk<C>(H, H));
    A[2][2] = mul_si<SIGN>(A[2][2]);
    A[2][3] = mul_tw<SIGN>(A[2][3], mk<C>(-H, H));
    A[3][1] = mul_tw<SIGN>(A[3][1], mk<C>(S1, C1));
    A[3][2] = mul_tw<SIGN>(A[3][2], mk<C>(-H, H));
    A[3][3] = mul_tw<SIGN>(A[3][3], mk<C>(-C1, -S1));
#pragma unroll
    for (int j1 = 0; j1 < 4; j1++) {
      C c[4] = {A[0][j1], A[1][j1], A[2][j1], A[3][j1]};
      Dft<4, SIGN>::run(c);
#pragma unroll
      for (int j2 = 0; j2 < 4; j2++) v[j1 + 4 * j2] = c[j2];
    }
  }
};

// Radix schedules (product = N). Larger radices first keep register use flat.
template <int N> struct Sched;
template <int... Rs> struct RList {};
template <> struct Sched<16> { using L = RList<16>; };
template <> struct Sched<24> { using L = RList<8, 3>; };
template <> struct Sched<32> { using L = RList<8, 4>; };
template <> struct Sched<48> { using L = RList<16, 3>; };
template <> struct Sched<64> { using L = RList<8, 8>; };
template <> struct Sched<96> { using L = RList<8, 4, 3>; };   // 16.6 (one exchange fewer): 19 % slower in pools
template <> struct Sched<128> { using L = RList<16, 8>; };
template <> struct Sched<192> { using L = RList<16, 4, 3>; };
template <> struct Sched<256> { using L = RList<16, 16>; };
template <> struct Sched<384> { using L = RList<16, 8, 3>; };
template <> struct Sched<512> { using L = RList<8, 8, 8>; };
template <> struct Sched<768> { using L = RList<16, 16, 3>; };
template <> struct Sched<1024> { using L = RList<16, 8, 8>; };

// LDS layouts. Interleaved (pencil passes, T a multiple of 8): element n of line t at n*T + t, threads walk t
// fastest so global loads and stores coalesce across lines. LineMajor (x pass, few lines): line t holds its N
// elements contiguously with one pad slot after every 8, and threads walk the butterfly index fastest, so the
// radix-stride writes of the first stage spread over the LDS banks.
template <int N, int T> struct Interleaved {
  static constexpr int SIZE = N * T;
  static __device__ __forceinline__ int at(int n, int t) { return n * T + t; }
  template <int NB> static __device__ __forceinline__ void split(int idx, int &t, int &bb) { t = idx % T; bb = idx / T; }
};
// SW selects the slot of element n within its line: 0, one pad slot after every 8 (per 4: x pass 184 -> 212 us at
// 256^3; per 16: 187); 1, none; 2, the low 4 bits of n XORed with bits 4..7; 3, the low 3 bits XORed with bits 3..5;
// 4, one pad slot after every 16.
template <int N, int T, int SW = 0> struct LineMajor {
  static_assert(SW == 0 || SW == 1 || (SW == 2 && N % 16 == 0) || (SW == 3 && N % 8 == 0) || (SW == 4 && N % 16 == 0),
                "a swizzle permutes whole blocks of the line");
  static constexpr int STRIDE = SW == 0 ? N + N / 8 : SW == 4 ? N + N / 16 : N, SIZE = T * STRIDE;
  static __device__ __forceinline__ int slot(int n) {
    if constexpr (SW == 0) return n + (n >> 3);
    else if constexpr (SW == 1) return n;
    else if constexpr (SW == 2) return n ^ ((n >> 4) & 15);
    else if constexpr (SW == 3) return n ^ ((n >> 3) & 7);
    else return n + (n >> 4);
  }
  static __device__ __forceinline__ int at(int n, int t) { return t * STRIDE + slot(n); }
  template <int NB> static __device__ __forceinline__ void split(int idx, int &t, int &bb) { bb = idx % NB; t = idx / NB; }
};

// The thread's index within the NT threads that run a transform: a one-wave transform (NT = 64) may run in each wave
// of a larger workgroup, every wave on lines of its own.
template <int NT> __device__ __forceinline__ int xf_tid() { return NT == 64 ? (int)(threadIdx.x & 63) : (int)threadIdx.x; }

// Twiddle table values of one thread for every stage of a radix schedule, read once. A kernel that runs several
// transforms in the same layout (the x pass: five per row) reads them before the first transform and keeps them in
// registers, so no stage waits on a table read. Slot j of stage S holds tw[j * p * S] for the j that twiddles<R>
// reads from the table; the other powers are formed from these exactly as twiddles<R> forms them.
__host__ __device__ constexpr bool tw_from_table(int j) { return !NSL_TW_PRODUCTS || j < 4 || (j & 3) == 0; }
template <int N, int T, int NT, class LY, int S, class L, class C = double2> struct TwPre {
  __device__ __forceinline__ void load(const C *__restrict__) {}
};
template <int N, int T, int NT, class LY, int S, int R0, int... Rs, class C> struct TwPre<N, T, NT, LY, S, RList<R0, Rs...>, C> {
  static constexpr int NB = N / R0, TOT = NB * T, K = (TOT + NT - 1) / NT;
  static constexpr bool HAS = S * R0 != N;
  C w[HAS ? K : 1][R0];
  TwPre<N, T, NT, LY, S * R0, RList<Rs...>, C> next;
  __device__ __forceinline__ void load(const C *__restrict__ tw) {
    if constexpr (HAS) {
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int idx = xf_tid<NT>() + k * NT;
        int t, bb;
        LY::template split<NB>(idx, t, bb);
        const int base = (bb / S) * S;
#pragma unroll
        for (int j = 1; j < R0; j++)
          if (tw_from_table(j)) w[k][j] = (TOT % NT == 0 || idx < TOT) ? tw[j * base] : mk<C>(1, 0);
      }
    }
    next.load(tw);
  }
  template <int SS> __device__ __forceinline__ auto &at() {
    if constexpr (SS == S) return *this;
    else return next.template at<SS>();
  }
};
template <int N, int T, int NT, class LY, class SCH, class C = double2> using TwCache = TwPre<N, T, NT, LY, 1, SCH, C>;

// The twiddles of slot k of a stage: from the table (TWS a pointer) or from a TwCache.
template <int R, int S, class TWS, class C>
__device__ __forceinline__ void stage_twiddles(TWS &tws, int k, int base, C (&w)[R]) {
  if constexpr (!std::is_class<typename std::remove_cv<TWS>::type>::value) {   // the table itself
    twiddles<R>(tws, base, w);
  } else {
    auto &st = tws.template at<S>();
#pragma unroll
    for (int j = 1; j < R; j++) {
      if (tw_from_table(j)) w[j] = st.w[k][j];
      else w[j] = cmul(w[j & ~3], w[j & 3]);
    }
  }
}

// Barrier between stages. LW = false: the workgroup's barrier. LW = true: an LDS-only barrier. Each wave completes
// its own LDS operations (s_waitcnt lgkmcnt(0)); a workgroup of more than one wave (NT > 64) then meets at s_barrier.
// An empty asm statement with a memory clobber on each side keeps the compiler from moving any memory operation
// across. A single-thread fence does not: lanes exchange data through LDS addresses that the compiler sees as
// distinct, and it reordered them. Unlike a workgroup fence, this does not wait for outstanding global loads,
// including loads into LDS still in flight.
template <bool LW, int NT = 64> __device__ __forceinline__ void stage_sync() {
  if constexpr (LW) {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0); vmcnt and expcnt not waited on
    if constexpr (NT > 64) __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  } else {
    __syncthreads();
  }
}

// One stage. S = product of the radices before it. Input either from `load(n, t)` (FIRST) or LDS; output to
// `store(n, t, v)` (LAST) or LDS, at the layout LY's addresses.
template <int N, int T, int NT, int R, int S, bool FIRST, bool LAST, int SIGN, bool SYNC_FIRST, bool LW, class LY,
          class TWS, class LD, class ST, class C>
__device__ __forceinline__ void stage(C *__restrict__ lds, TWS &tw, LD &load, ST &store) {
  constexpr int NB = N / R, TOT = NB * T, K = (TOT + NT - 1) / NT;
  C v[K][R];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int idx = xf_tid<NT>() + k * NT;
    if (TOT % NT == 0 || idx < TOT) {
      int t, bb;
      LY::template split<NB>(idx, t, bb);
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int n = bb + j * NB;
        if constexpr (!FIRST) v[k][j] = lds[LY::at(n, t)];
        // a load functor may also take the register slot (k, j): inputs can come from registers
        else if constexpr (std::is_invocable<LD, int, int, int, int>::value) v[k][j] = load(n, t, k, j);
        else v[k][j] = load(n, t);
      }
    }
  }
  if (!FIRST || SYNC_FIRST) stage_sync<LW, NT>();
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int idx = xf_tid<NT>() + k * NT;
    if (TOT % NT == 0 || idx < TOT) {
      int t, bb;
      LY::template split<NB>(idx, t, bb);
      const int p = bb / S, q = bb % S;
      Dft<R, SIGN>::run(v[k]);
      if (S * R != N) {
        C w[R];
        stage_twiddles<R, S>(tw, k, p * S, w);
#pragma unroll
        for (int j = 1; j < R; j++) v[k][j] = mul_tw<SIGN>(v[k][j], w[j]);
      }
      const int o0 = q + S * R * p;
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int n = o0 + j * S;
        if constexpr (LAST) {
          // a store functor may also take the register slot (k, j): outputs can stay in registers
          if constexpr (std::is_invocable<ST, int, int, C, int, int>::value) store(n, t, v[k][j], k, j);
          else store(n, t, v[k][j]);
        } else {
          lds[LY::at(n, t)] = v[k][j];
        }
      }
    }
  }
  if (!LAST) stage_sync<LW, NT>();
}

// Register footprint of the last stage: outputs land in slots (k, j) at element n = bb + j*(N/R), line t, where
// idx = threadIdx.x + k*NT, t = idx % T, bb = idx / T (Interleaved layout).
template <class L> struct LastR;
template <int R0> struct LastR<RList<R0>> { static constexpr int R = R0; };
template <int R0, int R1, int... Rs> struct LastR<RList<R0, R1, Rs...>> { static constexpr int R = LastR<RList<R1, Rs...>>::R; };
template <int N, int T, int NT, class C = double2> struct OutRegs {
  static constexpr int R = LastR<typename Sched<N>::L>::R, S = N / R, TOT = S * T, K = (TOT + NT - 1) / NT;
  C v[K][R];
  static __device__ __forceinline__ bool valid(int k) { return TOT % NT == 0 || (int)threadIdx.x + k * NT < TOT; }
  static __device__ __forceinline__ int line(int k) { return ((int)threadIdx.x + k * NT) % T; }
  static __device__ __forceinline__ int elem(int k, int j) { return ((int)threadIdx.x + k * NT) / T + j * S; }
};

template <int N, int T, int NT, int SIGN, bool SYNC_FIRST, class LY, int S, bool FIRST, bool LW = false, class TWS,
          class LD, class ST, class C, int R0, int... Rs>
__device__ __forceinline__ void run_stages(C *lds, TWS &tw, LD &load, ST &store, RList<R0, Rs...>) {
  constexpr bool LAST = sizeof...(Rs) == 0;
  stage<N, T, NT, R0, S, FIRST, LAST, SIGN, SYNC_FIRST, LW, LY>(lds, tw, load, store);
  if constexpr (!LAST)
    run_stages<N, T, NT, SIGN, SYNC_FIRST, LY, S * R0, false, LW>(lds, tw, load, store, RList<Rs...>{});
}

// Full transform of T lines of length N. tw: the N roots of unity (cos, sin)(2 pi k / N). SYNC_FIRST: the first
// stage's loads read LDS that the stage then overwrites (in-place use), so a barrier separates them.
template <int N, int T, int NT, int SIGN, bool SYNC_FIRST, class LY = Interleaved<N, T>, class TWS, class LD, class ST,
          class C>
__device__ __forceinline__ void fft(C *lds, TWS &tw, LD &load, ST &store) {
  run_stages<N, T, NT, SIGN, SYNC_FIRST, LY, 1, true>(lds, tw, load, store, typename Sched<N>::L{});
}
// The same with an explicit radix schedule SCH (an RList whose product is N); tw is the table or a
// TwCache<N, T, NT, LY, SCH>.
template <int N, int T, int NT, int SIGN, bool SYNC_FIRST, class LY, class SCH, class TWS, class LD, class ST, class C>
__device__ __forceinline__ void fft_sched(C *lds, TWS &tw, LD &load, ST &store) {
  run_stages<N, T, NT, SIGN, SYNC_FIRST, LY, 1, true>(lds, tw, load, store, SCH{});
}
// The same with LDS-only barriers (stage_sync<true, NT>), for kernels with loads into LDS in flight.
template <int N, int T, int NT, int SIGN, bool SYNC_FIRST, class LY, class SCH, class TWS, class LD, class ST, class C>
__device__ __forceinline__ void fft_sched_lds(C *lds, TWS &tw, LD &load, ST &store) {
  run_stages<N, T, NT, SIGN, SYNC_FIRST, LY, 1, true, true>(lds, tw, load, store, SCH{});
}

// First radix of a schedule (and the rest of it).
template <class L> struct FirstR;
template <int R0, int... Rs> struct FirstR<RList<R0, Rs...>> {
  static constexpr int R = R0;
  using Rest = RList<Rs...>;
  static constexpr bool single = sizeof...(Rs) == 0;
};

}  // namespace fast
}  // namespace nsl
