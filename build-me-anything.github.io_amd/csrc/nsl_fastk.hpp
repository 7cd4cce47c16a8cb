// Size-specialised pass kernels (P1..P5 of one rhs) over the compile-time FFT engine in nsl_fast.hpp.
// Same data layouts and arithmetic as the generic kernels in nsl_kernels.hpp; only the transform engine differs
// (radix-8/16 stages in registers, first stage read from HBM, last stage written to HBM).
#pragma once
#include <type_traits>

#include "nsl_fast.hpp"
#include "nsl_kernels.hpp"

namespace nsl {
namespace fast {

// The 2/3-truncated geometry as a compile-time constant of N (these kernels serve the solver state only).
template <int N>
__host__ __device__ constexpr Geo trunc_geo() {
  Geo g{};
  g.N = N; g.NH = N / 2 + 1; g.hx = N / 3 + 1; g.by = 2 * (N / 3) + 1; g.hy = N / 3 + 1;
  g.hp = (g.hx + 7) & ~7;
  g.D = (long long)g.by * g.by * g.hp;
  return g;
}

// Threads per workgroup of the pencil passes (P1, P2, P4, P5). With 8 lines of N >= 512 per workgroup, 512 threads
// give each thread one butterfly per stage, which halves the registers per thread and doubles the waves per CU.
#ifndef NSL_PNT_LARGE
#define NSL_PNT_LARGE 512
#endif
// (At 256^3 the 16 x 16 schedule has 128 butterflies per stage for 8 lines, so half of a 256-thread workgroup idles;
// 128-thread workgroups measured no faster: y-inverse 177.5-178 against 173-174 us, y-forward 83.5 against 82.8.)
template <int N> struct PassNT { static constexpr int v = N >= 512 ? NSL_PNT_LARGE : NSL_NT; };

// Pencil passes store whole tiles, the padding columns kx = hx .. hp-1 included (zeros), so each 128-byte line is
// written in full. FP64 tiles (T lines of 16 bytes) divide the row pitch hp; the y passes' FP32 tiles of 16 lines
// (128-byte rows of complex64) may run past it by 8 columns, which are neither read nor written (Past<N, T, E>::cut).
#ifndef NSL_FULL_LINES
#define NSL_FULL_LINES 1
#endif

// First-stage slot j of a pencil transform along y or z reads elements j*NB .. j*NB+NB-1 (NB = N / first radix).
// True when none of them is a retained wavenumber: the slot is zero and needs no load.
template <int N, int T, class E> struct Past {
  static constexpr bool cut = trunc_geo<N>().hp % T != 0;
  static_assert(!cut || !std::is_same<E, double2>::value, "FP64 tiles cover the row pitch exactly");
  static_assert(!cut || T == 16, "only 16-line FP32 tiles run past the row pitch");
};
template <int N> __device__ __forceinline__ constexpr bool zero_slot(int j) {
  constexpr Geo g = trunc_geo<N>();
  constexpr int NB = N / FirstR<typename Sched<N>::L>::R;
  return j * NB >= g.hy && j * NB + NB - 1 < N - g.by + g.hy;
}

// P1: X[o][z][cy][kx] = z-inverse of field o (MODE 0 stored, 1 u/omega, 2 gradient of gc). Grid ntile*nout*by,
// tile fastest: neighbouring workgroups write neighbouring segments of the same rows, and the nout workgroups that
// read the same state tile run close together (their re-reads of the state hit L2).
template <int N, int T, int MODE>
__global__ __launch_bounds__(PassNT<N>::v) void kf_zinv(const double2 *__restrict__ U, double2 *__restrict__ X, const Geo,
                                                   const double2 *__restrict__ tw, int nout, int gc,
                                                   const Ctl *__restrict__ ctl, size_t sb) {
  U = at_slot(U, sb); X = at_slot(X, sb); ctl = at_slot(ctl, sb);
  if (ctl != nullptr && ctl->active == 0) return;
  extern __shared__ double2 lds[];
  constexpr Geo g = trunc_geo<N>();
  const int ntile = (g.hx + T - 1) / T;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = bid % ntile, rest = bid / ntile;
  const int o = rest % nout, cy = rest / nout, kx0 = tile * T;
  const double ky = kwv(full_of(cy, g), N);
  const int plane = g.by * g.hp;
  const size_t ubase = (size_t)cy * g.hp + kx0;
  auto load = [&](int n, int t) -> double2 {
    const int kx = kx0 + t, cz = compact_of(n, g);
    if (cz < 0 || kx >= g.hx) return make_double2(0.0, 0.0);
    return zsrc<MODE>(U, g.D, ubase + (size_t)cz * plane + t, o, (double)kx, ky, kwv(n, N), gc);
  };
  double2 *Xb = X + ((size_t)o * N * g.by + cy) * g.hp + kx0;
  auto store = [&](int n, int t, double2 v) {
    if (kx0 + t < g.hx) Xb[(size_t)n * plane + t] = v;
  };
  fft<N, T, PassNT<N>::v, 1, false>(lds, tw, load, store);
}

// P1 for the rhs (u and omega), two workgroups per (kx tile, ky row). The curl is formed after the z transform,
// in (kx, ky, z) space, from A_c = IFFT_z(u_c) and B_c = IFFT_z(i kz u_c): omega-hat = i k x u-hat (nslab.js:203-212)
// is linear, so IFFT_z(omega-hat_x) = i ky A_z - B_y, IFFT_z(omega-hat_y) = B_x - i kx A_z,
// IFFT_z(omega-hat_z) = i (kx A_y - ky A_x). Type 0 transforms u_x, u_y and writes u_x, u_y, omega_z; type 1
// transforms u_z, i kz u_y, i kz u_x and writes u_z, omega_x, omega_y: five transforms and five state-tile reads per
// tile instead of six and nine. Grid ntile * 2 * by, XCD-remapped so the pair shares an L2.
// Waves per SIMD the z-inverse must fit: at N >= 512 (one butterfly per thread) a cap of 128 registers costs no
// spills and lets two workgroups share a CU.
// Above 512 the 8-line tile's LDS (96 KB at 768, 128 KB at 1024) allows one workgroup per CU whatever the register
// count, so the cap is 256 registers (no spills) instead of 128.
#ifndef NSL_ZINV_BIG_MINW
#define NSL_ZINV_BIG_MINW 2
#endif
template <int N> struct ZinvMinWaves { static constexpr int v = N > 512 ? NSL_ZINV_BIG_MINW : N == 512 ? 4 : 1; };
// Threads of the z-inverse: at N = 256 (16 x 16, 16 outputs per butterfly) 128 threads, one butterfly each.
template <int N> struct ZinvNT { static constexpr int v = N == 256 ? 128 : PassNT<N>::v; };
// nfin > 0: one workgroup more than the tiles need, the grid's first (dispatched first, so its single-workgroup
// sums overlap the tiles rather than trailing them), books the previous rhs if its fused update left that pending
// (fin_rhs: its nfin partials, diagnostics into stage fmode), in this launch instead of a kernel of its own; the
// tiles' workgroups never read what it writes. Its sums use the tile's LDS (N T complex) as scratch.
template <int N, int T> struct ZinvFin {
  static constexpr int SCR = N * T * 2 / NSL_NT;   // doubles per virtual thread the LDS holds
  static constexpr int CH = SCR >= 4 ? 4 : SCR >= 2 ? 2 : SCR >= 1 ? 1 : 0;
};
template <int N, int T, class E = double2>
__global__ __launch_bounds__(ZinvNT<N>::v, ZinvMinWaves<N>::v) void kf_zinv_uw(const E *__restrict__ U, E *__restrict__ X, const Geo,
                                                      const E *__restrict__ tw, Ctl *__restrict__ ctl, size_t sb,
                                                      const double *__restrict__ partials, int nfin, int fmode) {
  using C = E;   // the transforms and the curl run in the storage precision
  using F = Sc<C>;
  U = at_slot(U, sb); X = at_slot(X, sb); ctl = at_slot(ctl, sb);
  if (ctl != nullptr && ctl->active == 0) return;
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  const int first = nfin > 0 ? 1 : 0, ntiles = (int)gridDim.x - first;
  if ((int)blockIdx.x < first) {
    if constexpr (ZinvFin<N, T>::CH > 0)
      if (ctl->pend != 0)
        fin_rhs<ZinvNT<N>::v, ZinvFin<N, T>::CH>(at_slot(partials, sb), nfin, fmode, nullptr, ctl, (double *)lds_raw);
    return;
  }
  constexpr Geo g = trunc_geo<N>();
  static_assert(g.hp % T == 0, "whole tiles cover the row pitch");
  constexpr int ntile = (g.hx + T - 1) / T, plane = g.by * g.hp;
  using OR = OutRegs<N, T, ZinvNT<N>::v, C>;
  const int bid = xcd_remap((int)blockIdx.x - first, ntiles);
  const int tile = bid % ntile, rest = bid / ntile, type = rest & 1, cy = rest >> 1, kx0 = tile * T;
  const F ky = (F)kwv(full_of(cy, g), N);
  const size_t ubase = (size_t)cy * g.hp + kx0;
  E *Xb = X + (size_t)cy * g.hp + kx0;
  OR A, B;
  auto keepA = [&](int, int, C v, int k, int j) { A.v[k][j] = v; };
  auto keepB = [&](int, int, C v, int k, int j) { B.v[k][j] = v; };
  // Below N = 512 the i kz factor multiplies the selected value (zero off the retained band) after the load's
  // branch, so no branch holds arithmetic on a loaded value and a transform's loads stay in flight together
  // (256^3: 118 -> 110 us). At N >= 512 that needs more than the 128-register cap and spills (993 -> 1046 us), so
  // the product stays on the load's path there.
  auto load_c = [&](int c, bool ikz) {
    return [=](int n, int t, int, int j) -> C {
      const int cz = compact_of(n, g);
      if constexpr (N >= 512) {
        if (cz < 0 || kx0 + t >= g.hx) return mk<C>(0, 0);
        const C u = U[(size_t)c * g.D + ubase + (size_t)cz * plane + t];
        if (!ikz) return u;
        const F kz = (F)kwv(n, N);
        return mk<C>(-kz * u.y, kz * u.x);
      } else {
        if (zero_slot<N>(j)) return mk<C>(0, 0);
        const C u = (cz >= 0 && kx0 + t < g.hx) ? U[(size_t)c * g.D + ubase + (size_t)cz * plane + t] : mk<C>(0, 0);
        const F kz = (F)kwv(n, N);
        return ikz ? mk<C>(-kz * u.y, kz * u.x) : u;
      }
    };
  };
  // out(o, f): write f(k, j) to field o at every register slot of this thread
  auto out = [&](int o, auto f) {   // whole tiles, padding columns (zero inputs, zero outputs) included
#pragma unroll
    for (int k = 0; k < OR::K; k++) {
      if (!OR::valid(k) || (!NSL_FULL_LINES && kx0 + OR::line(k) >= g.hx)) continue;
#pragma unroll
      for (int j = 0; j < OR::R; j++) Xb[((size_t)o * N + OR::elem(k, j)) * plane + OR::line(k)] = f(k, j);
    }
  };
  if (type == 0) {
    auto lx = load_c(0, false);
    fft<N, T, ZinvNT<N>::v, 1, false>(lds, tw, lx, keepA);
    out(0, [&](int k, int j) { return A.v[k][j]; });
    auto ly = load_c(1, false);
    fft<N, T, ZinvNT<N>::v, 1, false>(lds, tw, ly, keepB);
    out(1, [&](int k, int j) { return B.v[k][j]; });
    out(5, [&](int k, int j) {
      const F kx = (F)(kx0 + OR::line(k));
      const F pr = kx * B.v[k][j].x - ky * A.v[k][j].x, pi = kx * B.v[k][j].y - ky * A.v[k][j].y;
      return mk<C>(-pi, pr);
    });
  } else {
    auto lz = load_c(2, false);
    fft<N, T, ZinvNT<N>::v, 1, false>(lds, tw, lz, keepA);
    out(2, [&](int k, int j) { return A.v[k][j]; });
    auto ly = load_c(1, true);
    fft<N, T, ZinvNT<N>::v, 1, false>(lds, tw, ly, keepB);
    out(3, [&](int k, int j) {
      return mk<C>(-ky * A.v[k][j].y - B.v[k][j].x, ky * A.v[k][j].x - B.v[k][j].y);
    });
    auto lx = load_c(0, true);
    fft<N, T, ZinvNT<N>::v, 1, false>(lds, tw, lx, keepB);
    out(4, [&](int k, int j) {
      const F kx = (F)(kx0 + OR::line(k));
      return mk<C>(B.v[k][j].x + kx * A.v[k][j].y, B.v[k][j].y - kx * A.v[k][j].x);
    });
  }
}

// P2: X1[o][z][cy][kx] -> X2[o][z][y][kx]. Grid (ntile*N, ncomp).
template <int N, int T, class E = double2>
__global__ __launch_bounds__(PassNT<N>::v) void kf_yinv(const E *__restrict__ X1, E *__restrict__ X2, const Geo,
                                                   const E *__restrict__ tw, const Ctl *__restrict__ ctl, size_t sb) {
  using C = E;   // the transform runs in the storage precision
  X1 = at_slot(X1, sb); X2 = at_slot(X2, sb); ctl = at_slot(ctl, sb);
  if (ctl != nullptr && ctl->active == 0) return;
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  constexpr Geo g = trunc_geo<N>();
  constexpr bool cut = Past<N, T, E>::cut;
  const int ntile = (g.hx + T - 1) / T;
  const int tile = blockIdx.x % ntile, z = blockIdx.x / ntile, o = blockIdx.y, kx0 = tile * T;
  const E *Ib = X1 + ((size_t)o * N + z) * g.by * g.hp + kx0;
  E *Ob = X2 + ((size_t)o * N + z) * N * g.hp + kx0;
  auto load = [&](int n, int t, int, int j) -> C {
    if (zero_slot<N>(j)) return mk<C>(0, 0);
    const int cy = compact_of(n, g);
    return (cy >= 0 && kx0 + t < g.hx) ? Ib[cy * g.hp + t] : mk<C>(0, 0);
  };
  // every column of the tile is stored, padding kx >= hx included (its inputs are zero, so it stores zeros): whole
  // 128-byte lines, never a partial line
  auto store = [&](int n, int t, C v) {
    if ((NSL_FULL_LINES || kx0 + t < g.hx) && (!cut || kx0 + t < g.hp)) Ob[n * g.hp + t] = v;
  };
  fft<N, T, PassNT<N>::v, 1, false>(lds, tw, load, store);
}

// P4: X3[c][z][y][kx] -> X4[c][z][cy][kx] (retained ky). Grid (ntile*N, ncomp).
template <int N, int T, class E = double2>
__global__ __launch_bounds__(PassNT<N>::v) void kf_yfwd(const E *__restrict__ X3, E *__restrict__ X4, const Geo,
                                                   const E *__restrict__ tw, const Ctl *__restrict__ ctl, size_t sb,
                                                   const double *__restrict__ pmax, int nblk, int nred) {
  using C = E;   // the transform runs in the storage precision
  X3 = at_slot(X3, sb); X4 = at_slot(X4, sb); ctl = at_slot(ctl, sb);
  if (ctl != nullptr && ctl->active == 0) return;
  // The x pass's per-workgroup maxima (max |omega|^2, max |u|^2), when given: the first nred workgroups of component 0
  // combine them into the control block, in place of a reduction kernel of their own. The x pass has completed before
  // this pass starts, and the update that reads the maxima starts after it.
  if (pmax != nullptr && ctl != nullptr && blockIdx.y == 0 && (int)blockIdx.x < nred) {
    constexpr int NT = PassNT<N>::v;
    __shared__ double mscr[2 * (NT / 64)];
    pmax = at_slot(pmax, sb);
    double a = 0.0, b = 0.0;
    for (int i = blockIdx.x * NT + threadIdx.x; i < nblk; i += nred * NT) {
      a = fmax(a, pmax[2 * (size_t)i]);
      b = fmax(b, pmax[2 * (size_t)i + 1]);
    }
    Ctl *cw = const_cast<Ctl *>(ctl);
    block_atomic_max2<NT>(a, b, &cw->om2, &cw->u2, mscr);
  }
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  constexpr Geo g = trunc_geo<N>();
  constexpr bool cut = Past<N, T, E>::cut;
  const int ntile = (g.hx + T - 1) / T;
  const int tile = blockIdx.x % ntile, z = blockIdx.x / ntile, c = blockIdx.y, kx0 = tile * T;
  const E *Ib = X3 + ((size_t)c * N + z) * N * g.hp + kx0;
  E *Ob = X4 + ((size_t)c * N + z) * g.by * g.hp + kx0;
  auto load = [&](int n, int t) -> C {
    return kx0 + t < g.hx ? Ib[n * g.hp + t] : mk<C>(0, 0);
  };
  auto store = [&](int n, int t, C v) {   // whole tiles, as in the y-inverse
    const int cy = compact_of(n, g);
    if (cy >= 0 && (NSL_FULL_LINES || kx0 + t < g.hx) && (!cut || kx0 + t < g.hp)) Ob[cy * g.hp + t] = v;
  };
  fft<N, T, PassNT<N>::v, -1, false>(lds, tw, load, store);
}

// P5: X4[c][z][cy][kx] -> OUT[c][cz][cy][kx] * inv (retained kz). Grid (ntile*by, ncomp).
template <int N, int T>
__global__ __launch_bounds__(PassNT<N>::v) void kf_zfwd(const double2 *__restrict__ X4, double2 *__restrict__ OUT, const Geo,
                                                   const double2 *__restrict__ tw, double inv, const Ctl *__restrict__ ctl,
                                                   size_t sb) {
  X4 = at_slot(X4, sb); OUT = at_slot(OUT, sb); ctl = at_slot(ctl, sb);
  if (ctl != nullptr && ctl->active == 0) return;
  extern __shared__ double2 lds[];
  constexpr Geo g = trunc_geo<N>();
  static_assert(g.hp % T == 0, "whole tiles cover the row pitch");
  const int ntile = (g.hx + T - 1) / T;
  const int tile = blockIdx.x % ntile, cy = blockIdx.x / ntile, c = blockIdx.y, kx0 = tile * T;
  const int plane = g.by * g.hp;
  const double2 *Ib = X4 + (size_t)c * N * plane + (size_t)cy * g.hp + kx0;
  double2 *Ob = OUT + (size_t)c * g.D + (size_t)cy * g.hp + kx0;
  auto load = [&](int n, int t) -> double2 {
    return kx0 + t < g.hx ? Ib[(size_t)n * plane + t] : make_double2(0.0, 0.0);
  };
  auto store = [&](int n, int t, double2 v) {   // whole tiles, as in the y-inverse
    const int cz = compact_of(n, g);
    if (cz >= 0 && (NSL_FULL_LINES || kx0 + t < g.hx)) Ob[(size_t)cz * plane + t] = make_double2(v.x * inv, v.y * inv);
  };
  fft<N, T, PassNT<N>::v, -1, false>(lds, tw, load, store);
}

// P5 fused with the spectral update: one workgroup transforms all three components of its T lines (kx0..kx0+T-1
// at row cy) along z, keeps the retained outputs in registers and applies update_mode<STAGE> to each of them, so the
// nonlinear term never goes to HBM. The radix schedule ends in radix 8 and a thread owns one last-stage butterfly
// per line: slot j of thread x is element n = x/T + j*N/8 of line x%T, for all three components at once.
// Sums: one partial per workgroup (fixed-order tree), grid ntile*by.
template <int N> struct ZuSched;
template <> struct ZuSched<32> { using L = RList<4, 8>; };
template <> struct ZuSched<48> { using L = RList<2, 3, 8>; };   // 6.8: no difference in pools or alone
template <> struct ZuSched<64> { using L = RList<8, 8>; };
template <> struct ZuSched<96> { using L = RList<4, 3, 8>; };
template <> struct ZuSched<128> { using L = RList<16, 8>; };
template <> struct ZuSched<192> { using L = RList<8, 3, 8>; };
template <> struct ZuSched<256> { using L = RList<8, 4, 8>; };
template <> struct ZuSched<384> { using L = RList<16, 3, 8>; };
template <> struct ZuSched<512> { using L = RList<8, 8, 8>; };
template <> struct ZuSched<1024> { using L = RList<16, 8, 8>; };
// Alternative schedules the autotuner may pick instead (variant 1; same last radix 8, so the same thread-to-mode
// ownership): more, smaller stages at the front.
template <int N, int V> struct ZuSchedV { using L = typename ZuSched<N>::L; };
template <> struct ZuSchedV<32, 1> { using L = RList<2, 2, 8>; };
template <> struct ZuSchedV<48, 1> { using L = RList<6, 8>; };
template <> struct ZuSchedV<64, 1> { using L = RList<4, 2, 8>; };
template <> struct ZuSchedV<96, 1> { using L = RList<2, 6, 8>; };
template <> struct ZuSchedV<128, 1> { using L = RList<4, 4, 8>; };
template <> struct ZuSchedV<192, 1> { using L = RList<4, 6, 8>; };
template <> struct ZuSchedV<256, 1> { using L = RList<4, 8, 8>; };
template <> struct ZuSchedV<384, 1> { using L = RList<8, 6, 8>; };
template <> struct ZuSchedV<512, 1> { using L = RList<16, 4, 8>; };
template <> struct ZuSchedV<1024, 1> { using L = RList<8, 16, 8>; };
#ifndef NSL_ZU_T
#define NSL_ZU_T 8
#endif
#ifndef NSL_ZU_GROUP
#define NSL_ZU_GROUP 1
#endif
// Lines per workgroup: 8 (whole 128-byte rows); 16 at 32^3 (one wave: 4 threads per line, and the row pitch is 16);
// 4 at 1024^3, where 8 lines would need 1024 threads with at most 128 registers each.
// (24 lines at 48^3 and 64^3, the whole row pitch, measured 2 % slower in pools)
__host__ __device__ constexpr int zu_lines(int N) { return N > 512 ? 4 : N == 32 ? 16 : NSL_ZU_T; }
template <int N> struct ZuCfg {
  static constexpr int T = zu_lines(N), R = 8, NB = N / R, NT = NB * T;
  // at most 256 registers (512 from 512 threads up: one workgroup per CU, no accumulation registers needed);
  // a cap of 168 or 128 spills 58-146 registers
  static constexpr int MINW = NT >= 512 ? 1 : 2;
};
// The last-stage slots j (elements j*N/8 .. j*N/8 + N/8 - 1) that hold a retained kz, in order.
template <int N> struct ZuSlots {
  static constexpr Geo g = trunc_geo<N>();
  static constexpr bool has(int j) { return !(j * (N / 8) >= g.hy && j * (N / 8) + N / 8 - 1 < N - g.by + g.hy); }
  static constexpr int count() {
    int c = 0;
    for (int j = 0; j < 8; j++) c += has(j) ? 1 : 0;
    return c;
  }
  static constexpr int at(int i) {
    for (int j = 0, c = 0; j < 8; j++)
      if (has(j)) {
        if (c == i) return j;
        c++;
      }
    return 0;
  }
};
template <int N, int STAGE, int V = 0, class E = double2>
__global__ __launch_bounds__(ZuCfg<N>::NT, ZuCfg<N>::MINW) void kf_zfwd_upd(const E *__restrict__ X4, const E *U,
                                                            E *S, E *__restrict__ ACC, E *TS,
                                                            const Geo, const E *__restrict__ tw, double inv,
                                                            double *__restrict__ partials, Ctl *__restrict__ ctl,
                                                            size_t sb) {
  using C = ZuCfg<N>;
  using CT = E;   // the z-forward runs in the storage precision; the update below in FP64
  constexpr int T = C::T, R = C::R, NB = C::NB, NT = C::NT;
  X4 = at_slot(X4, sb); U = at_slot(U, sb); S = at_slot(S, sb); ACC = at_slot(ACC, sb); TS = at_slot(TS, sb);
  partials = at_slot(partials, sb); ctl = at_slot(ctl, sb);
  if (ctl->active == 0) return;
  extern __shared__ double2 lds_raw[];
  CT *lds = reinterpret_cast<CT *>(lds_raw);
  constexpr Geo g = trunc_geo<N>();
  static_assert(g.hp % T == 0, "whole tiles cover the row pitch");
  constexpr int ntile = (g.hx + T - 1) / T, plane = g.by * g.hp;
  // A tile narrower than a 128-byte line (4 lines at 1024^3, every FP32 tile) shares each line with its neighbour
  // tile: dealt to the XCDs in contiguous logical ranges, the two halves of a line are read through one L2 (1024^3:
  // update 13.5 -> 12.0 ms). Whole-line tiles keep the plain order (0.5-2 % faster at 256^3 and 512^3). The partial
  // sums are indexed by the logical tile, so the summation order is the same either way.
  const int bid = T * (int)sizeof(E) < 128 ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int tile = bid % ntile, cy = bid / ntile, kx0 = tile * T;
  const E *Ib = X4 + (size_t)cy * g.hp + kx0;
  double2 Nl[3][R];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    auto load = [&](int n, int t) -> CT {
      return kx0 + t < g.hx ? Ib[((size_t)c * N + n) * plane + t] : mk<CT>(0, 0);
    };
    auto keep = [&](int, int, CT v, int, int j) { Nl[c][j] = make_double2((double)v.x * inv, (double)v.y * inv); };
    fft_sched<N, T, NT, -1, false, Interleaved<N, T>, typename ZuSchedV<N, V>::L>(lds, tw, load, keep);
  }
  const double nu = ctl->nu;
  const double dt = STAGE == 1 ? choose_dt(ctl, sqrt(__longlong_as_double((long long)ctl->u2))) : ctl->dt;
  const int t = threadIdx.x % T, bb = threadIdx.x / T, ix = kx0 + t;
  const double kx = (double)ix, ky = kwv(full_of(cy, g), N);
  const double h = (ix == 0 || 2 * ix == N) ? 1.0 : 2.0;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // Tnl Pspec E Z Pal nE nZ nP
  // slots whose element range holds a retained kz, in groups of ZG: a group's loads are unconditional (a slot
  // with no retained kz, or a padding column kx >= hx, reads mode cz = 0 of its column and discards it) so they
  // issue back to back; only the writes and sums are conditional
  constexpr int ZG = NSL_ZU_GROUP, nj = ZuSlots<N>::count();
#pragma unroll
  for (int g0 = 0; g0 < nj; g0 += ZG) {
    UpdIn in[ZG];
#pragma unroll
    for (int q = 0; q < ZG; q++) {
      if (g0 + q >= nj) break;
      const int n = bb + ZuSlots<N>::at(g0 + q) * NB, cz = max(compact_of(n, g), 0);
      in[q] = upd_load<STAGE>(((size_t)cz * g.by + cy) * g.hp + ix, g.D, U, S, ACC);
    }
#pragma unroll
    for (int q = 0; q < ZG; q++) {
      if (g0 + q >= nj) break;
      const int j = ZuSlots<N>::at(g0 + q), n = bb + j * NB, cz = compact_of(n, g);
      if (cz < 0) continue;
      const size_t m = ((size_t)cz * g.by + cy) * g.hp + ix;
      if (ix >= g.hx) {   // padding column: zeros, so every 128-byte line of the outputs is written whole
        if (NSL_FULL_LINES) upd_zero<STAGE>(m, g.D, S, ACC, TS);
        continue;
      }
      upd_apply<STAGE>(in[q], m, g.D, kx, ky, kwv(n, N), h, Nl[0][j], Nl[1][j], Nl[2][j], S, ACC, TS, nu, dt, acc);
    }
  }
  block_sum<8, NT>(acc, reinterpret_cast<double *>(lds_raw), partials + (size_t)bid * 8);
  // the booking of these partials (fin_rhs) is left to the next launch that reads them: the next rhs's z-inverse,
  // the next step's k_begin_step, or a k_finalize at the end of the steps enqueued
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->pend = 1;
}

// LDS layout of the x pass (interleaved measured faster than line-major on gfx950).
#ifndef NSL_XNL_LINEMAJOR
#define NSL_XNL_LINEMAJOR 0
#endif
template <int N, int T>
using XnlLayout = typename std::conditional<NSL_XNL_LINEMAJOR != 0, LineMajor<N, T>, Interleaved<N, T>>::type;

// P3: in place on X2 (6 fields in, 3 out), LP row pairs per workgroup. Same packing and product as k_xnl.
template <int N, int LP, class E = double2>
__global__ __launch_bounds__(NSL_NT) void kf_xnl(E *__restrict__ X2, const Geo, const E *__restrict__ tw,
                                                  double *__restrict__ pmax, Ctl *__restrict__ ctl, size_t sb) {
  using C = E;   // the transforms and the product run in the storage precision
  using F = Sc<C>;
  X2 = at_slot(X2, sb); pmax = at_slot(pmax, sb); ctl = at_slot(ctl, sb);
  if (ctl->active == 0) return;
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  constexpr Geo g = trunc_geo<N>();
  constexpr int T = 6 * LP, T2 = 3 * LP, half = N / 2, npairs = N * half;
  using L1 = XnlLayout<N, T>;
  using L2 = XnlLayout<N, T2>;
  // retained kx modes and row pitch are functions of N (the 2/3 rule); compile-time so the loops below unroll and
  // every load of the workgroup is in flight at once
  constexpr int hx = N / 3 + 1, HP = (hx + 7) & ~7;
  constexpr int ITEMS = T * hx, KL = (ITEMS + NSL_NT - 1) / NSL_NT;
  constexpr size_t cs = (size_t)N * N * HP;
  constexpr int ngroups = (npairs + LP - 1) / LP;
  // one group of LP row pairs per workgroup (a persistent loop with prefetch needs 256+ VGPRs on gfx950 and
  // measured slower)
  C A[KL], B[KL];
  auto load_group = [&](int grp) {
#pragma unroll
    for (int i = 0; i < KL; i++) {
      const int e = threadIdx.x + i * NSL_NT;
      A[i] = B[i] = mk<C>(0.0, 0.0);
      if (ITEMS % NSL_NT == 0 || e < ITEMS) {
        const int t = e / hx, k = e - t * hx, p = t / 6, o = t - p * 6;
        const int pr = grp * LP + p;
        if (pr < npairs) {
          const int z = pr / half, y = 2 * (pr - z * half);
          const size_t base = o * cs + ((size_t)z * N + y) * HP;
          A[i] = (X2[base + k]);
          B[i] = (X2[base + HP + k]);
        }
      }
    }
  };
  double om2 = 0.0, u2 = 0.0;
  const int grp = blockIdx.x;
  load_group(grp);
  {
#pragma unroll
  for (int i = 0; i < KL; i++) {
    const int e = threadIdx.x + i * NSL_NT;
    if (ITEMS % NSL_NT == 0 || e < ITEMS) {
      const int t = e / hx, k = e - t * hx;
      lds[L1::at(k, t)] = mk<C>(A[i].x - B[i].y, A[i].y + B[i].x);
      if (k > 0 && 2 * k < N) lds[L1::at(N - k, t)] = mk<C>(A[i].x + B[i].y, B[i].x - A[i].y);
    }
  }
  __syncthreads();
  // modes hx .. N-hx of every line are zero: never written, never read
  auto ldl = [&](int n, int t) -> C {
    return (n >= hx && n <= N - hx) ? mk<C>(0.0, 0.0) : lds[L1::at(n, t)];
  };
  auto stl = [&](int n, int t, C v) { lds[L1::at(n, t)] = v; };
  fft<N, T, NSL_NT, 1, true, L1>(lds, tw, ldl, stl);
  __syncthreads();
  constexpr int KI = (N * LP + NSL_NT - 1) / NSL_NT;
  F nv[KI][6];
#pragma unroll
  for (int k2 = 0; k2 < KI; k2++) {
    const int it = threadIdx.x + k2 * NSL_NT;
    if ((N * LP) % NSL_NT == 0 || it < N * LP) {
      const int x = it % N, p = it / N;
      const C Za = lds[L1::at(x, p * 6)], Zb = lds[L1::at(x, p * 6 + 1)], Zc = lds[L1::at(x, p * 6 + 2)];
      const C Zp = lds[L1::at(x, p * 6 + 3)], Zq = lds[L1::at(x, p * 6 + 4)], Zr = lds[L1::at(x, p * 6 + 5)];
      {
        const F a = Za.x, b = Zb.x, c = Zc.x, pp = Zp.x, q = Zq.x, r = Zr.x;
        const F o2 = pp * pp + q * q + r * r, uu = a * a + b * b + c * c;
        om2 = fmax(om2, (double)o2); u2 = fmax(u2, (double)uu);
        nv[k2][0] = b * r - c * q; nv[k2][1] = c * pp - a * r; nv[k2][2] = a * q - b * pp;
      }
      {
        const F a = Za.y, b = Zb.y, c = Zc.y, pp = Zp.y, q = Zq.y, r = Zr.y;
        const F o2 = pp * pp + q * q + r * r, uu = a * a + b * b + c * c;
        om2 = fmax(om2, (double)o2); u2 = fmax(u2, (double)uu);
        nv[k2][3] = b * r - c * q; nv[k2][4] = c * pp - a * r; nv[k2][5] = a * q - b * pp;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < KI; k2++) {
    const int it = threadIdx.x + k2 * NSL_NT;
    if ((N * LP) % NSL_NT == 0 || it < N * LP) {
      const int x = it % N, p = it / N;
      lds[L2::at(x, p * 3)] = mk<C>(nv[k2][0], nv[k2][3]);
      lds[L2::at(x, p * 3 + 1)] = mk<C>(nv[k2][1], nv[k2][4]);
      lds[L2::at(x, p * 3 + 2)] = mk<C>(nv[k2][2], nv[k2][5]);
    }
  }
  __syncthreads();
  auto ldl2 = [&](int n, int t) -> C { return lds[L2::at(n, t)]; };
  // only modes k < hx and their mirrors N-k are unpacked
  auto stl2 = [&](int n, int t, C v) {
    if (n < hx || n > N - hx) lds[L2::at(n, t)] = v;
  };
  fft<N, T2, NSL_NT, -1, true, L2>(lds, tw, ldl2, stl2);
  __syncthreads();
  // whole rows, padding columns hx..HP-1 as zeros (every 128-byte line written in full)
  constexpr int ITEMS2 = T2 * HP, KS = (ITEMS2 + NSL_NT - 1) / NSL_NT;
#pragma unroll
  for (int i = 0; i < KS; i++) {
    const int e = threadIdx.x + i * NSL_NT;
    if (!(ITEMS2 % NSL_NT == 0 || e < ITEMS2)) continue;
    const int t = e / HP, k = e - t * HP, p = t / 3, c = t - p * 3;
    const int pr = grp * LP + p;
    if (pr >= npairs) continue;
    const int z = pr / half, y = 2 * (pr - z * half);
    C X = mk<C>(0.0, 0.0), Y = X;
    if (k < hx) {
      const C Zk = lds[L2::at(k, t)], Zc = lds[L2::at((N - k) % N, t)];
      const F zr = Zk.x, zi = Zk.y, cr = Zc.x, ci = -Zc.y;
      X = mk<C>(0.5 * (zr + cr), 0.5 * (zi + ci));
      Y = mk<C>(0.5 * (zi - ci), -0.5 * (zr - cr));
    }
    const size_t base = c * cs + ((size_t)z * N + y) * HP;
    X2[base + k] = X;
    X2[base + HP + k] = Y;
  }
  }
  block_max2_store(om2, u2, pmax + 2 * (size_t)blockIdx.x, (double *)lds_raw);
}

// P3 with RP rows per workgroup of NT threads (half the LDS of the row-pair form for RP = 1). Each row's six
// half-spectra ride as three complex lines, (u_x + i u_y), (u_z + i om_x), (om_y + i om_z); after the inverse the
// product u x omega is formed per point, and (n_x + i n_y), (n_z + i 0) go through the forward transform.
template <int N, int RP, int NT, class E = double2>
__global__ __launch_bounds__(NT) void kf_xnl_rows(E *__restrict__ X2, const Geo, const E *__restrict__ tw,
                                                   double *__restrict__ pmax, Ctl *__restrict__ ctl, size_t sb) {
  using C = E;   // the transforms and the product run in the storage precision
  using F = Sc<C>;
  X2 = at_slot(X2, sb); pmax = at_slot(pmax, sb); ctl = at_slot(ctl, sb);
  if (ctl->active == 0) return;
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  constexpr Geo g = trunc_geo<N>();
  constexpr int T = 3 * RP, T2 = 2 * RP, nrows = N * N;
  using L1 = Interleaved<N, T>;
  using L2 = Interleaved<N, T2>;
  constexpr int hx = g.hx, HP = g.hp;
  constexpr int ITEMS = T * hx, KL = (ITEMS + NT - 1) / NT;
  constexpr size_t cs = (size_t)N * N * HP;
  const int row0 = blockIdx.x * RP;
  C A[KL], B[KL];
#pragma unroll
  for (int i = 0; i < KL; i++) {
    const int e = threadIdx.x + i * NT;
    A[i] = B[i] = mk<C>(0.0, 0.0);
    if (ITEMS % NT == 0 || e < ITEMS) {
      const int t = e / hx, k = e - t * hx, r = t / 3, l = t - r * 3;   // line l of row r: fields 2l, 2l+1
      if (row0 + r < nrows) {
        const size_t base = (size_t)(row0 + r) * HP + k;
        A[i] = (X2[2 * l * cs + base]);
        B[i] = (X2[(2 * l + 1) * cs + base]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < KL; i++) {
    const int e = threadIdx.x + i * NT;
    if (ITEMS % NT == 0 || e < ITEMS) {
      const int t = e / hx, k = e - t * hx;
      lds[L1::at(k, t)] = mk<C>(A[i].x - B[i].y, A[i].y + B[i].x);
      if (k > 0 && 2 * k < N) lds[L1::at(N - k, t)] = mk<C>(A[i].x + B[i].y, B[i].x - A[i].y);
    }
  }
  __syncthreads();
  auto ldl = [&](int n, int t) -> C {
    return (n >= hx && n <= N - hx) ? mk<C>(0.0, 0.0) : lds[L1::at(n, t)];
  };
  auto stl = [&](int n, int t, C v) { lds[L1::at(n, t)] = v; };
  fft<N, T, NT, 1, true, L1>(lds, tw, ldl, stl);
  __syncthreads();
  constexpr int KI = (N * RP + NT - 1) / NT;
  F nv[KI][3];
  double om2 = 0.0, u2 = 0.0;
#pragma unroll
  for (int k2 = 0; k2 < KI; k2++) {
    const int it = threadIdx.x + k2 * NT;
    if ((N * RP) % NT == 0 || it < N * RP) {
      const int x = it % N, r = it / N;
      const C Z0 = lds[L1::at(x, 3 * r)], Z1 = lds[L1::at(x, 3 * r + 1)], Z2 = lds[L1::at(x, 3 * r + 2)];
      const F a = Z0.x, b = Z0.y, c = Z1.x, pp = Z1.y, q = Z2.x, rr = Z2.y;   // u = (a, b, c), omega = (pp, q, rr)
      const F o2 = pp * pp + q * q + rr * rr, uu = a * a + b * b + c * c;
      om2 = fmax(om2, (double)o2); u2 = fmax(u2, (double)uu);
      nv[k2][0] = b * rr - c * q; nv[k2][1] = c * pp - a * rr; nv[k2][2] = a * q - b * pp;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < KI; k2++) {
    const int it = threadIdx.x + k2 * NT;
    if ((N * RP) % NT == 0 || it < N * RP) {
      const int x = it % N, r = it / N;
      lds[L2::at(x, 2 * r)] = mk<C>(nv[k2][0], nv[k2][1]);
      lds[L2::at(x, 2 * r + 1)] = mk<C>(nv[k2][2], 0.0);
    }
  }
  __syncthreads();
  auto ldl2 = [&](int n, int t) -> C { return lds[L2::at(n, t)]; };
  auto stl2 = [&](int n, int t, C v) {
    if (n < hx || n > N - hx) lds[L2::at(n, t)] = v;
  };
  fft<N, T2, NT, -1, true, L2>(lds, tw, ldl2, stl2);
  __syncthreads();
  // whole rows, padding columns hx..HP-1 as zeros (every 128-byte line written in full)
  constexpr int ITEMS2 = RP * HP, KS = (ITEMS2 + NT - 1) / NT;
#pragma unroll
  for (int i = 0; i < KS; i++) {
    const int e = threadIdx.x + i * NT;
    if (!(ITEMS2 % NT == 0 || e < ITEMS2)) continue;
    const int r = e / HP, k = e - r * HP;
    if (row0 + r >= nrows) continue;
    const size_t base = (size_t)(row0 + r) * HP + k;
    const int kk = (N - k) % N;
    C o0 = mk<C>(0.0, 0.0), o1 = o0, o2 = o0;
    if (k < hx) {
      {  // n_x + i n_y
        const C Zk = lds[L2::at(k, 2 * r)], Zc = lds[L2::at(kk, 2 * r)];
        const F zr = Zk.x, zi = Zk.y, cr = Zc.x, ci = -Zc.y;
        o0 = mk<C>(0.5 * (zr + cr), 0.5 * (zi + ci));
        o1 = mk<C>(0.5 * (zi - ci), -0.5 * (zr - cr));
      }
      {  // n_z (imaginary part zero)
        const C Zk = lds[L2::at(k, 2 * r + 1)], Zc = lds[L2::at(kk, 2 * r + 1)];
        o2 = mk<C>(0.5 * (Zk.x + Zc.x), 0.5 * (Zk.y - Zc.y));
      }
    }
    X2[base] = o0;
    X2[cs + base] = o1;
    X2[2 * cs + base] = o2;
  }
  // per-workgroup maxima
  om2 = wave_max(om2); u2 = wave_max(u2);
  __syncthreads();
  double *scratch = (double *)lds_raw;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { scratch[2 * wid] = om2; scratch[2 * wid + 1] = u2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ma = scratch[0], mb = scratch[1];
    for (int w = 1; w < NT / 64; w++) { ma = fmax(ma, scratch[2 * w]); mb = fmax(mb, scratch[2 * w + 1]); }
    pmax[2 * (size_t)blockIdx.x] = ma; pmax[2 * (size_t)blockIdx.x + 1] = mb;
  }
}

// P3 with one wave per x row (or per two rows at N = 256). The radix schedule starts and ends with radix 8, so a
// line has N/8 first- and last-stage butterflies: 64 at N = 512 (one row per wave, 8 x 8 x 8) and 32 at N = 256 (two
// rows per wave, 8 x 4 x 8). Every lane owns butterfly bb of all three packed lines (u_x + i u_y), (u_z + i om_x),
// (om_y + i om_z) of its row, so the product u x omega is formed in registers from the inverse's last stage (outputs
// x = bb + (N/8) j) and the forward's first stage starts from those registers (inputs at the same x): LDS carries
// only the exchanges inside each line's transform, one line of each row at a time, and the half-spectra are read
// straight from HBM into the first stage (Hermitian mirror of modes N-k in place).
template <int N> struct XnlWave;
template <> struct XnlWave<32> { using L = RList<4, 2, 4>; };
template <> struct XnlWave<48> { using L = RList<4, 3, 4>; };
template <> struct XnlWave<64> { using L = RList<8, 8>; };
// 96^3: 4.6.4 (radix-6 middle stage; 4.3.2.4 took one LDS exchange more: 29.3 against 26.2 us alone)
template <> struct XnlWave<96> { using L = RList<4, 6, 4>; };
template <> struct XnlWave<128> { using L = RList<8, 2, 8>; };
// 192^3 and 384^3 with radix-4 ends (4.3.4.4, 4.4.3.4; 4.3.8.4, 4.8.3.4) need fewer registers (166-208) but add an
// LDS exchange: 150-205 us against 140-146 us at 192^3, 1646-1652 against 1477 us (8.3.2.8) at 384^3
template <> struct XnlWave<192> { using L = RList<8, 3, 8>; };
// 384^3: 8.6.8 (8.3.2.8: 1468 against 1192 us)
template <> struct XnlWave<384> { using L = RList<8, 6, 8>; };
template <> struct XnlWave<512> { using L = RList<8, 8, 8>; };
template <> struct XnlWave<256> { using L = RList<8, 4, 8>; };
template <int N, class E = double2>
__global__ __launch_bounds__(64) void kf_xnl_wave(E *__restrict__ X2, const Geo, const E *__restrict__ tw,
                                                  double *__restrict__ pmax, Ctl *__restrict__ ctl, size_t sb) {
  using C = E;   // the transforms and the product run in the storage precision
  using F = Sc<C>;
  constexpr int R = 8, NB = N / R, RPW = 64 / NB;   // butterflies per line and stage ends, rows per wave
  static_assert(NB * RPW == 64, "first and last stage must cover a wave");
  using SCH = typename XnlWave<N>::L;
  static_assert(FirstR<SCH>::R == R && LastR<SCH>::R == R, "schedule starts and ends in radix 8");
  X2 = at_slot(X2, sb); pmax = at_slot(pmax, sb); ctl = at_slot(ctl, sb);
  if (ctl->active == 0) return;
  extern __shared__ double2 lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  constexpr Geo g = trunc_geo<N>();
  constexpr int hx = g.hx, HP = g.hp;
  constexpr size_t cs = (size_t)N * N * HP;
  using LY = LineMajor<N, RPW>;
  const int lane = threadIdx.x;
  const int row0 = blockIdx.x * RPW;
  // the twiddles of this lane's butterflies are the same in all five transforms: read once
  TwCache<N, RPW, 64, LY, SCH, C> twc;
  twc.load(tw);
  C U[3][R];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    // line l of row row0 + t: mode n < hx from the half-spectrum, n > N - hx from its mirror, zero between. Every
    // load is unconditional (the zero band reads index 0 and discards it), so a line's loads issue back to back and
    // are waited on once; slots j whose range j*NB .. j*NB+NB-1 lies in the zero band load nothing.
    auto ld = [&](int n, int t, int, int j) -> C {
      if (j * NB >= hx && j * NB + NB - 1 <= N - hx) return mk<C>(0.0, 0.0);
      const size_t base = (size_t)(row0 + t) * HP;
      const E *Fa = X2 + 2 * l * cs + base, *Fb = X2 + (2 * l + 1) * cs + base;
      const bool lo = n < hx, hi = n > N - hx;
      const int m = lo ? n : (hi ? N - n : 0);
      const C A = Fa[m], B = Fb[m];
      const C d = lo ? mk<C>(A.x - B.y, A.y + B.x) : mk<C>(A.x + B.y, B.x - A.y);
      return lo || hi ? d : mk<C>(0.0, 0.0);
    };
    auto keep = [&](int, int, C v, int, int j) { U[l][j] = v; };
    fft_sched<N, RPW, 64, 1, false, LY, SCH>(lds, twc, ld, keep);
  }
  // u x omega at x = bb + NB j of row row0 + lane / NB: the inverse's outputs of this lane
  double om2 = 0.0, u2 = 0.0;
  C P[2][R];
#pragma unroll
  for (int j = 0; j < R; j++) {
    const F a = U[0][j].x, b = U[0][j].y, c = U[1][j].x, pp = U[1][j].y, q = U[2][j].x, rr = U[2][j].y;
    om2 = fmax(om2, (double)(pp * pp + q * q + rr * rr)); u2 = fmax(u2, (double)(a * a + b * b + c * c));
    P[0][j] = mk<C>(b * rr - c * q, c * pp - a * rr);
    P[1][j] = mk<C>(a * q - b * pp, 0.0);
  }
#pragma unroll
  for (int l = 0; l < 2; l++) {
    auto ld2 = [&](int, int, int, int j) -> C { return P[l][j]; };
    auto st2 = [&](int n, int t, C v) {   // only modes k < hx and their mirrors N-k are unpacked
      if (n < hx || n > N - hx) lds[LY::at(n, t)] = v;
    };
    fft_sched<N, RPW, 64, -1, false, LY, SCH>(lds, twc, ld2, st2);
    __syncthreads();
    // whole rows, padding columns hx..HP-1 included (as zeros): every 128-byte line is written in full, which
    // the memory system retires far faster than a partial line (1474 -> 1269 us for this access pattern alone)
#pragma unroll
    for (int i = 0; i < (RPW * HP + 63) / 64; i++) {
      const int e = lane + i * 64;
      if (e >= RPW * HP) continue;
      const int t = e / HP, k = e - t * HP;
      const size_t base = (size_t)(row0 + t) * HP;
      C o0 = mk<C>(0.0, 0.0), o1 = o0;
      if (k < hx) {
        const C Zk = lds[LY::at(k, t)], Zc = lds[LY::at((N - k) % N, t)];
        if (l == 0) {   // n_x + i n_y
          const F zr = Zk.x, zi = Zk.y, cr = Zc.x, ci = -Zc.y;
          o0 = mk<C>(0.5 * (zr + cr), 0.5 * (zi + ci));
          o1 = mk<C>(0.5 * (zi - ci), -0.5 * (zr - cr));
        } else {        // n_z (imaginary part zero)
          o0 = mk<C>(0.5 * (Zk.x + Zc.x), 0.5 * (Zk.y - Zc.y));
        }
      }
      if (l == 0) {
        X2[base + k] = o0;
        X2[cs + base + k] = o1;
      } else {
        X2[2 * cs + base + k] = o0;
      }
    }
    __syncthreads();   // the next line's first stage overwrites LDS
  }
  om2 = wave_max(om2); u2 = wave_max(u2);
  if (lane == 0) { pmax[2 * (size_t)blockIdx.x] = om2; pmax[2 * (size_t)blockIdx.x + 1] = u2; }
}

// The x pass with one butterfly per thread per stage and its half-spectra staged through LDS by the load-to-LDS
// instruction: line l+1's input is in flight while line l is transformed, and a persistent workgroup stages its next
// row group's first line during the current one's last transforms. Rows per workgroup RPW and threads NT: 2 and 64
// at 256^3 (8x4x8), 1 and 64 at 512^3 (8x8x8), 1 and 128 at 1024^3 (8x16x8; two waves per row). The transforms use
// LDS-only barriers (stage_sync<true, NT>), since a workgroup fence would also wait for the staging loads in flight.
// Same arithmetic per line as kf_xnl_wave.
typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;
// 1024^3: 8x8x2x8 with table twiddles, 248 registers (two waves per SIMD), 17.9 ms against 18.5 ms for 8x16x8 with
// cached twiddles (329 registers, one wave per SIMD). Two waves per row; 4x8x8x4 and 4x4x4x4x4 with four waves per
// row, and staging whole rows instead of lines, measured no faster.
template <> struct XnlWave<1024> { using L = RList<8, 8, 2, 8>; };
#ifndef NSL_XNL_BIG_MINW
#define NSL_XNL_BIG_MINW 1
#endif
__host__ __device__ constexpr int gcd_int(int a, int b) { return b == 0 ? a : gcd_int(b, a % b); }
// The staged form's shape: stage-end radix R (8, or 4 where N has no 8.....8 factorisation: 32, 48, 96), NB = N/R
// butterflies per line, and the smallest workgroup of whole waves that holds whole rows, lcm(NB, 64) threads: one
// wave for RPW = 64/NB rows up to 512^3, two waves per row at 1024^3, three waves for 8 or 16 rows at 48, 96, 192
// and 384.
// Alternative schedules of the staged x pass the autotuner may pick (variant 1; same first and last radix, which
// sets the workgroup's shape: NT = lcm(N/R, 64) threads).
template <int N, int V> struct XnlSched { using L = typename XnlWave<N>::L; };
template <> struct XnlSched<32, 1> { using L = RList<2, 8, 2>; };
template <> struct XnlSched<64, 1> { using L = RList<4, 4, 4>; };
template <> struct XnlSched<96, 1> { using L = RList<4, 3, 2, 4>; };
template <> struct XnlSched<128, 1> { using L = RList<4, 8, 4>; };
template <> struct XnlSched<192, 1> { using L = RList<4, 3, 4, 4>; };
template <> struct XnlSched<256, 1> { using L = RList<4, 16, 4>; };
template <> struct XnlSched<384, 1> { using L = RList<8, 3, 2, 8>; };
template <> struct XnlSched<512, 1> { using L = RList<8, 4, 2, 8>; };
template <> struct XnlSched<1024, 1> { using L = RList<8, 16, 8>; };
// Variant 2 (N = 48, 96, 192, 384: N/R does not divide 64): the default schedule in one wave per row group, RPW rows
// of N/R butterflies with the remaining lanes idle in the first and last stages (the largest RPW <= 64/(N/R) that
// divides N^2), so no stage needs a workgroup barrier and the twiddles stay in registers.
template <int N> struct XnlOneWave {
  static constexpr int NB = N / FirstR<typename XnlWave<N>::L>::R;
  static constexpr bool ok = NB < 64 && 64 % NB != 0;
  static constexpr int rpw(int r) { return r <= 1 || (N * N) % r == 0 ? r : rpw(r - 1); }
  static constexpr int RPW = ok ? rpw(64 / NB) : 1;
};
template <int N, int V = 0> struct XnlStaged {
  static constexpr int R = FirstR<typename XnlSched<N, V>::L>::R, NB = N / R;
  static constexpr bool ONE = V == 2;
  static_assert(!ONE || XnlOneWave<N>::ok, "variant 2 is the one-wave shape of a multi-wave size");
  static constexpr int NT = ONE ? 64 : NB / gcd_int(NB, 64) * 64, RPW = ONE ? XnlOneWave<N>::RPW : NT / NB;
  static constexpr int WPR = NT / 64, ACT = NB * RPW;   // ACT: threads with a first- and last-stage butterfly
  // two-wave rows (1024^3): at most 512 / NSL_XNL_BIG_MINW registers; three-wave groups: at most 256
  // (three waves per SIMD, 168 registers, spills 95 at 192^3 and 111 at 384^3)
  static constexpr int MINW = NT == 128 ? NSL_XNL_BIG_MINW : NT > 64 ? 2 : 1;
  // one-wave groups hold the twiddles of all five transforms in registers (TwCache); multi-wave groups read them at
  // each stage from a copy of the table entries their schedule uses, made in LDS when the workgroup starts
  // (TwLen): at 1024^3 (8.8.2.8) a register cache is the difference between 248 and 300+ registers (two waves per
  // SIMD or one), at 192^3 it spills 14 registers. Reading the global table instead costs a vector-memory wait in
  // every stage, and vmcnt retires in issue order, so each such wait also waited for the staging loads of the next
  // line in flight: the 1024^3 x pass took 18.1 ms that way, 16.9 ms with the LDS copy (and 228 instead of 248
  // registers; the 384^3 staged form no longer spills).
  static constexpr bool TWC = NT == 64;
};
// Table entries a radix schedule reads (largest index + 1): stage (R, S) reads tw[j * base] for the powers j that
// twiddles<R> takes from the table and base = (bb / S) * S, bb < N / R (the last stage reads none).
template <int N, int S, class L> struct TwLen { static constexpr int v = 1; };
template <int N, int S, int R0, int... Rs> struct TwLen<N, S, RList<R0, Rs...>> {
  static constexpr int jmax() {
    int m = 0;
    for (int j = 1; j < R0; j++)
      if (tw_from_table(j)) m = j;
    return m;
  }
  static constexpr int mine = S * R0 == N ? 1 : jmax() * (((N / R0 - 1) / S) * S) + 1;
  static constexpr int rest = TwLen<N, S * R0, RList<Rs...>>::v;
  static constexpr int v = mine > rest ? mine : rest;
};
// The staged x pass's LDS layout (LineMajor's SW), from the x pass timed alone on every layout, each launch on a
// fresh copy of the same buffers: the low 3 bits XORed with bits 3..5 (SW 3) is the fastest or tied at every size
// (256^3 178 -> 171 us, 192^3 120 -> 117 us, 384^3 1046 -> 1021 us, 128^3 28.9 -> 27.8 us, 1024^3 17.2 -> 16.7 ms,
// 512^3 equal); the 4-bit XOR (SW 2) and one pad slot per 16 (SW 4) fall between, and no padding (SW 1) is 20-30 %
// slower everywhere.
template <int N> __host__ __device__ constexpr int xnl_sw() { return N % 8 == 0 ? 3 : 0; }
// One-wave groups unpack the forward transforms' outputs straight from the last stage's registers, the partner mode
// N - k taken across the wave (ds_bpermute), instead of a round trip through LDS (bit-identical; timed alone: 256^3
// 176.6 -> 167.9 us, 512^3 1379 -> 1352, 192^3 107.9 -> 101.3, 384^3 727 -> 680, 128^3 29.2 -> 27.9, 64^3 9.8 -> 9.3).
#ifndef NSL_XNL_DIRECT
#define NSL_XNL_DIRECT 1
#endif
template <int N, int V = 0, int SW = xnl_sw<N>(), bool DO = NSL_XNL_DIRECT != 0>
__global__ __launch_bounds__((XnlStaged<N, V>::NT), (XnlStaged<N, V>::MINW)) void kf_xnl_dma(double2 *__restrict__ X2, const Geo,
                                                               const double2 *__restrict__ tw,
                                                               double *__restrict__ pmax, Ctl *__restrict__ ctl,
                                                               size_t sb) {
  using C = XnlStaged<N, V>;
  constexpr int R = C::R, NB = C::NB, NT = C::NT, RPW = C::RPW, WPR = C::WPR, ACT = C::ACT;
  static_assert(ACT == NT || (C::ONE && ACT < NT), "first and last stage: one butterfly per thread (or idle lanes)");
  using SCH = typename XnlSched<N, V>::L;
  // the forward transforms start from the inverse's last-stage registers: slot j of the radix-R last stage holds
  // element bb + j*N/R, which is what a radix-R first stage reads, so the schedule starts and ends in radix R
  static_assert(FirstR<SCH>::R == R && LastR<SCH>::R == R, "schedule starts and ends in the same radix");
  constexpr Geo g = trunc_geo<N>();
  constexpr int hx = g.hx, HP = g.hp;
  constexpr size_t cs = (size_t)N * N * HP;
  using LY = LineMajor<N, RPW, SW>;
  constexpr int CH = (RPW * HP + 63) / 64, FB = CH * 64;   // 1 KB chunks per field of the RPW rows
  __shared__ double2 stg[2 * FB];
  __shared__ double2 xch[LY::SIZE];
  X2 = at_slot(X2, sb); pmax = at_slot(pmax, sb); ctl = at_slot(ctl, sb);
  if (ctl->active == 0) return;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // rows r .. r+RPW-1 are contiguous in each field: element e of the block is row t = e / HP, column e % HP. Wave w
  // stages chunks w, w + WPR, ...; lane i of chunk c lands at stg[c*64 + i].
  auto stage_line = [&](int l, int r) {
    const double2 *fa = X2 + 2 * l * cs + (size_t)r * HP, *fb = X2 + (2 * l + 1) * cs + (size_t)r * HP;
#pragma unroll
    for (int c = 0; c < CH; c++) {
      const int e = c * 64 + lane;
      if (c % WPR == wave && e < RPW * HP) {
        __builtin_amdgcn_global_load_lds((glb_void_t *)(fa + e), (lds_void_t *)(stg + c * 64), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((glb_void_t *)(fb + e), (lds_void_t *)(stg + FB + c * 64), 16, 0, 0);
      }
    }
  };
  // multi-wave groups: the table entries of the schedule in LDS (copies, so the same values as the table)
  constexpr int TLEN = C::TWC ? 1 : TwLen<N, 1, SCH>::v;
  static_assert(TLEN <= N, "twiddle indices stay inside the table");
  __shared__ double2 twl[TLEN];
  if constexpr (!C::TWC) {
    for (int i = tid; i < TLEN; i += NT) twl[i] = tw[i];
    __syncthreads();
  }
  stage_line(0, blockIdx.x * RPW);
  using TWC = typename std::conditional<C::TWC, TwCache<N, RPW, NT, LY, SCH>, const double2 *>::type;
  TWC twc;
  if constexpr (C::TWC) twc.load(tw);
  else twc = twl;
  // first stage: radix R0, K0 butterflies per thread (slot (k, j) is element bb + j*NB0 of line t)
  using F0 = FirstR<SCH>;
  constexpr int R0 = F0::R, NB0 = N / R0, TOT0 = NB0 * RPW, K0 = (TOT0 + NT - 1) / NT;
  static_assert(K0 * NT == TOT0 || C::ONE, "the first stage covers the workgroup");
  int t0[K0], bb0[K0];
#pragma unroll
  for (int k = 0; k < K0; k++) LY::template split<NB0>(tid + k * NT, t0[k], bb0[k]);
  const bool act = !C::ONE || tid < ACT;   // this thread owns a last-stage butterfly (all threads unless C::ONE)
  double om2 = 0.0, u2 = 0.0;
  for (int row0 = blockIdx.x * RPW; row0 < N * N; row0 += gridDim.x * RPW) {
  const int next = row0 + gridDim.x * RPW;
  double2 U[3][R];
  if constexpr (C::ONE) {   // idle lanes keep defined values (their products feed no store and no maximum)
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int j = 0; j < R; j++) U[l][j] = make_double2(0.0, 0.0);
  }
#pragma unroll
  for (int l = 0; l < 3; l++) {
    // line l's staging loads have landed: vmcnt(0) in every wave (explicit: the compiler's own wait for loads into
    // LDS was seen placed after the first read of the staged data), then the workgroup meets
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if constexpr (WPR > 1) __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    double2 v[K0][R0];
#pragma unroll
    for (int k = 0; k < K0; k++)
#pragma unroll
      for (int j = 0; j < R0; j++) {
        v[k][j] = make_double2(0.0, 0.0);
        // elements of slot (k, j): bb0 + j*NB0 with bb0 in [k*NT, k*NT + NT) for one row per workgroup, in
        // [0, NB0) otherwise; a slot wholly in the zero band reads nothing
        constexpr int span = RPW == 1 ? (NT < NB0 ? NT : NB0) : NB0;
        const int lo_n = j * NB0 + (RPW == 1 ? k * NT : 0);
        if (lo_n >= hx && lo_n + span - 1 <= N - hx) continue;
        if (TOT0 % NT != 0 && tid + k * NT >= TOT0) continue;   // idle lane (one-wave shape)
        const int n = bb0[k] + j * NB0;
        const bool lo = n < hx, hi = n > N - hx;
        const int m = t0[k] * HP + (lo ? n : (hi ? N - n : 0));
        const double2 A = stg[m], B = stg[FB + m];
        const double2 d = lo ? make_double2(A.x - B.y, A.y + B.x) : make_double2(A.x + B.y, B.x - A.y);
        v[k][j] = lo || hi ? d : make_double2(0.0, 0.0);
      }
    stage_sync<true, NT>();       // every staging read is complete before the next line overwrites stg
    if (l < 2) stage_line(l + 1, row0);
    else if (next < N * N) stage_line(0, next);
    auto ldr = [&](int, int, int k, int j) -> double2 { return v[k][j]; };
    auto keep = [&](int, int, double2 w, int, int j) { U[l][j] = w; };
    fft_sched_lds<N, RPW, NT, 1, false, LY, SCH>(xch, twc, ldr, keep);
  }
  double2 P[2][R];
#pragma unroll
  for (int j = 0; j < R; j++) {
    const double a = U[0][j].x, b = U[0][j].y, c = U[1][j].x, pp = U[1][j].y, q = U[2][j].x, rr = U[2][j].y;
    if (act) { om2 = fmax(om2, pp * pp + q * q + rr * rr); u2 = fmax(u2, a * a + b * b + c * c); }
    P[0][j] = make_double2(b * rr - c * q, c * pp - a * rr);
    P[1][j] = make_double2(a * q - b * pp, 0.0);
  }
#pragma unroll
  for (int l = 0; l < 2; l++) {
    auto ld2 = [&](int, int, int, int j) -> double2 { return P[l][j]; };
    if constexpr (DO && WPR == 1) {
      // the unpacking straight from the last stage's registers: slot j of this lane holds mode k = bb + j NB of row t,
      // and its partner N - k sits in lane t NB + (NB - bb) % NB, slot R - 1 - j (bb != 0), or in this lane, slot
      // (R - j) % R (bb == 0); the partner's value comes across the wave (ds_bpermute), the stores go out per slot
      double2 Z[R];
      if constexpr (C::ONE) {   // idle lanes keep defined values (no active lane reads them)
#pragma unroll
        for (int j = 0; j < R; j++) Z[j] = make_double2(0.0, 0.0);
      }
      auto keepZ = [&](int, int, double2 w, int, int j) { Z[j] = w; };
      fft_sched_lds<N, RPW, NT, -1, false, LY, SCH>(xch, twc, ld2, keepZ);
      const int bbL = lane % NB, tL = lane / NB, src = (tL * NB + (NB - bbL) % NB) * 4;
      auto bperm = [&](double v) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_ds_bpermute(src, (int)(unsigned)(b & 0xffffffffll));
        const int hi = __builtin_amdgcn_ds_bpermute(src, (int)(unsigned)((unsigned long long)b >> 32));
        return __longlong_as_double(((long long)(unsigned)hi << 32) | (long long)(unsigned)lo);
      };
      constexpr int JOUT = (HP + NB - 1) / NB;   // slots holding a stored column k < HP
      const size_t base = (size_t)(row0 + tL) * HP;
#pragma unroll
      for (int j = 0; j < JOUT; j++) {
        const double2 w = Z[R - 1 - j];
        double2 Zc = make_double2(bperm(w.x), bperm(w.y));
        if (bbL == 0) Zc = Z[(R - j) % R];
        const int k = bbL + j * NB;
        if (!act || k >= HP) continue;
        const double2 Zk = Z[j];
        double2 o0 = make_double2(0.0, 0.0), o1 = o0;
        if (k < hx) {
          if (l == 0) {
            const double zr = Zk.x, zi = Zk.y, cr = Zc.x, ci = -Zc.y;
            o0 = make_double2(0.5 * (zr + cr), 0.5 * (zi + ci));
            o1 = make_double2(0.5 * (zi - ci), -0.5 * (zr - cr));
          } else {
            o0 = make_double2(0.5 * (Zk.x + Zc.x), 0.5 * (Zk.y - Zc.y));
          }
        }
        if (l == 0) {
          stc(X2[base + k], o0);
          stc(X2[cs + base + k], o1);
        } else {
          stc(X2[2 * cs + base + k], o0);
        }
      }
      continue;
    }
    auto st2 = [&](int n, int t, double2 w) {
      if (n < hx || n > N - hx) xch[LY::at(n, t)] = w;
    };
    fft_sched_lds<N, RPW, NT, -1, false, LY, SCH>(xch, twc, ld2, st2);
    stage_sync<true, NT>();
    // whole rows, padding columns hx..HP-1 as zeros (every 128-byte line written in full)
#pragma unroll
    for (int i = 0; i < (RPW * HP + NT - 1) / NT; i++) {
      const int e = tid + i * NT;
      if (e >= RPW * HP) continue;
      const int t = e / HP, k = e - t * HP;
      const size_t base = (size_t)(row0 + t) * HP;
      double2 o0 = make_double2(0.0, 0.0), o1 = o0;
      if (k < hx) {
        const double2 Zk = xch[LY::at(k, t)], Zc = xch[LY::at((N - k) % N, t)];
        if (l == 0) {
          const double zr = Zk.x, zi = Zk.y, cr = Zc.x, ci = -Zc.y;
          o0 = make_double2(0.5 * (zr + cr), 0.5 * (zi + ci));
          o1 = make_double2(0.5 * (zi - ci), -0.5 * (zr - cr));
        } else {
          o0 = make_double2(0.5 * (Zk.x + Zc.x), 0.5 * (Zk.y - Zc.y));
        }
      }
      if (l == 0) {
        stc(X2[base + k], o0);
        stc(X2[cs + base + k], o1);
      } else {
        stc(X2[2 * cs + base + k], o0);
      }
    }
    stage_sync<true, NT>();   // the next line's first stage overwrites xch
  }
  }
  // one (max |omega|^2, max |u|^2) pair per workgroup, pmax[2 blockIdx.x]: the buffer holds 2 N^2 doubles and the
  // grid is at most N^2 / RPW workgroups (the maxima are exact, so the combining order does not matter)
  om2 = wave_max(om2); u2 = wave_max(u2);
  if constexpr (WPR > 1) {
    __shared__ double red[2 * WPR];
    if (lane == 0) { red[2 * wave] = om2; red[2 * wave + 1] = u2; }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < WPR; w++) { om2 = fmax(om2, red[2 * w]); u2 = fmax(u2, red[2 * w + 1]); }
      pmax[2 * (size_t)blockIdx.x] = om2; pmax[2 * (size_t)blockIdx.x + 1] = u2;
    }
  } else {
    if (lane == 0) { pmax[2 * (size_t)blockIdx.x] = om2; pmax[2 * (size_t)blockIdx.x + 1] = u2; }
  }
}

// ------------------------------------------------------------------------------------------------ plane pass
// P2-P4 of one z-plane in one workgroup, for grids whose six y-transformed fields fit the 160 KB of LDS (N = 32:
// 48 KB, N = 48: 108 KB, N = 64: 135 KB in rows of the retained columns): the plane's six fields come from X1 and are
// y-inverted in place in LDS, the x pass runs on
// the plane's rows there (one wave per group of RPW rows, the staged x pass's shape and arithmetic), and the three
// products are y-transformed forward into X4. X2 (6 h N^2 complex values written, 9 h N^2 read and 3 h N^2 written
// again by the three passes) never goes to HBM. X4 is X1's first three fields: each workgroup writes only the plane
// it read. Maxima: one atomic pair per workgroup into the control block, as the y-forward pass combines them.
// Lines of NF fields of one plane in LDS, field-major: line t = field t / HP, column t % HP; element n of a line at
// field * N * HP + n * HP + column, so a field's rows are the x pass's rows.
template <int N, int HP, int NF> struct PlaneLY {
  static constexpr int T = NF * HP, SIZE = NF * N * HP;
  static __device__ __forceinline__ int at(int n, int t) { return (t / HP) * N * HP + n * HP + t % HP; }
  template <int NB> static __device__ __forceinline__ void split(int idx, int &t, int &bb) { t = idx % T; bb = idx / T; }
};
template <int N> struct PlaneCfg {
  static constexpr Geo g = trunc_geo<N>();
  // 48^3: 512 threads (one workgroup per CU, limited by LDS: eight waves, 12 row groups of the x pass). 64^3: rows of
  // the hx retained columns only (135 KB for the plane, and three waves' exchange buffers in the rest of the 160 KB),
  // 192 threads, and the y transforms one field at a time (one butterfly per thread and stage).
  static constexpr int HP = N == 64 ? g.hx : g.hp, NT = N == 48 ? 512 : N == 64 ? 192 : 256, NW = NT / 64;
  static constexpr bool PER_FIELD = N == 64;
  static constexpr int XV = XnlOneWave<N>::ok ? 2 : 0;   // the staged x pass's one-wave shape
  using XC = XnlStaged<N, XV>;
  static_assert(XC::NT == 64, "one wave per row group");
  static constexpr int RPW = XC::RPW;
  using LX = LineMajor<N, RPW, xnl_sw<N>()>;
  static_assert(N % RPW == 0, "whole row groups per plane");
  static constexpr int PLANE = 6 * N * HP, XCH = NW * LX::SIZE;
  static constexpr int LDS_BYTES = (PLANE + XCH) * (int)sizeof(double2);
};
static constexpr bool plane_has(int N) { return N == 32 || N == 48 || N == 64; }

template <int N>
__global__ __launch_bounds__(PlaneCfg<N>::NT) void kf_plane(double2 *X1, const Geo, const double2 *__restrict__ tw,
                                                Ctl *__restrict__ ctl, size_t sb) {
  using C = PlaneCfg<N>;
  using XC = typename C::XC;
  constexpr Geo g = trunc_geo<N>();
  constexpr int HP = C::HP, NT = C::NT, RPW = C::RPW, R = XC::R, NB = XC::NB, hx = g.hx;
  static_assert(C::LDS_BYTES <= 160 * 1024, "the plane fits the LDS");
  X1 = at_slot(X1, sb); ctl = at_slot(ctl, sb);
  if (ctl->active == 0) return;
  extern __shared__ double2 lds_raw[];
  double2 *P = lds_raw;                                          // [field][y][kx], 6 fields
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, z = blockIdx.x;
  double2 *xch = lds_raw + C::PLANE + wave * C::LX::SIZE;       // this wave's exchange buffer
  const size_t plane = (size_t)g.by * g.hp;
  // 1. y-inverse of the six fields: X1[o][z][cy][kx] -> P[o][y][kx] (zero band and padding columns read as zeros)
  if constexpr (!C::PER_FIELD) {
    using LY = PlaneLY<N, HP, 6>;
    auto load = [&](int n, int t, int, int j) -> double2 {
      if (zero_slot<N>(j)) return make_double2(0.0, 0.0);
      const int cy = compact_of(n, g), o = t / HP, kx = t % HP;
      return (cy >= 0 && kx < hx) ? X1[((size_t)o * N + z) * plane + (size_t)cy * g.hp + kx] : make_double2(0.0, 0.0);
    };
    auto store = [&](int n, int t, double2 v) { P[LY::at(n, t)] = v; };
    fft<N, LY::T, NT, 1, false, LY>(P, tw, load, store);
  } else {
    using LY = PlaneLY<N, HP, 1>;
#pragma unroll 1
    for (int o = 0; o < 6; o++) {   // each field's transform writes only its own rows of P
      auto load = [&](int n, int t, int, int j) -> double2 {
        if (zero_slot<N>(j)) return make_double2(0.0, 0.0);
        const int cy = compact_of(n, g);
        return (cy >= 0 && t < hx) ? X1[((size_t)o * N + z) * plane + (size_t)cy * g.hp + t] : make_double2(0.0, 0.0);
      };
      auto store = [&](int n, int t, double2 v) { P[o * N * HP + LY::at(n, t)] = v; };
      fft<N, LY::T, NT, 1, false, LY>(P + o * N * HP, tw, load, store);
    }
  }
  __syncthreads();
  // 2. the x pass on the plane's rows (kf_xnl_dma's packing, product and unpacking; its inputs and outputs in P)
  using LY = typename C::LX;
  using SCH = typename XnlWave<N>::L;
  TwCache<N, RPW, 64, LY, SCH> twc;
  twc.load(tw);
  using F0 = FirstR<SCH>;
  constexpr int R0 = F0::R, NB0 = N / R0, TOT0 = NB0 * RPW, K0 = (TOT0 + 63) / 64;
  static_assert(R0 == R && K0 == 1, "one first-stage butterfly per lane");
  int t0, bb0;
  LY::template split<NB0>(lane, t0, bb0);
  const bool act = lane < NB * RPW;
  double om2 = 0.0, u2 = 0.0;
  for (int row0 = wave * RPW; row0 < N; row0 += C::NW * RPW) {
    double2 U[3][R];
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int j = 0; j < R; j++) U[l][j] = make_double2(0.0, 0.0);
#pragma unroll
    for (int l = 0; l < 3; l++) {
      double2 v[R0];
#pragma unroll
      for (int j = 0; j < R0; j++) {
        v[j] = make_double2(0.0, 0.0);
        const int lo_n = j * NB0;
        if (lo_n >= hx && lo_n + NB0 - 1 <= N - hx) continue;
        if (lane >= TOT0) continue;
        const int n = bb0 + j * NB0;
        const bool lo = n < hx, hi = n > N - hx;
        const int m = (row0 + t0) * HP + (lo ? n : (hi ? N - n : 0));
        const double2 A = P[(2 * l) * N * HP + m], B = P[(2 * l + 1) * N * HP + m];
        const double2 d = lo ? make_double2(A.x - B.y, A.y + B.x) : make_double2(A.x + B.y, B.x - A.y);
        v[j] = lo || hi ? d : make_double2(0.0, 0.0);
      }
      auto ldr = [&](int, int, int, int j) -> double2 { return v[j]; };
      auto keep = [&](int, int, double2 w, int, int j) { U[l][j] = w; };
      fft_sched_lds<N, RPW, 64, 1, false, LY, SCH>(xch, twc, ldr, keep);
    }
    double2 Pr[2][R];
#pragma unroll
    for (int j = 0; j < R; j++) {
      const double a = U[0][j].x, b = U[0][j].y, c = U[1][j].x, pp = U[1][j].y, q = U[2][j].x, rr = U[2][j].y;
      if (act) { om2 = fmax(om2, pp * pp + q * q + rr * rr); u2 = fmax(u2, a * a + b * b + c * c); }
      Pr[0][j] = make_double2(b * rr - c * q, c * pp - a * rr);
      Pr[1][j] = make_double2(a * q - b * pp, 0.0);
    }
    // the wave's reads of its rows in P are complete (each stage waits for its LDS operations), so the products
    // overwrite fields 0..2 of the same rows
    // unpacking straight from the last stage's registers, the partner mode N - k across the wave (as in kf_xnl_dma)
    const int bbL = lane % NB, tL = lane / NB, src = (tL * NB + (NB - bbL) % NB) * 4;
    auto bperm = [&](double v) {
      const long long b = __double_as_longlong(v);
      const int lo = __builtin_amdgcn_ds_bpermute(src, (int)(unsigned)(b & 0xffffffffll));
      const int hi = __builtin_amdgcn_ds_bpermute(src, (int)(unsigned)((unsigned long long)b >> 32));
      return __longlong_as_double(((long long)(unsigned)hi << 32) | (long long)(unsigned)lo);
    };
#pragma unroll
    for (int l = 0; l < 2; l++) {
      auto ld2 = [&](int, int, int, int j) -> double2 { return Pr[l][j]; };
      double2 Z[R];
#pragma unroll
      for (int j = 0; j < R; j++) Z[j] = make_double2(0.0, 0.0);
      auto keepZ = [&](int, int, double2 w, int, int j) { Z[j] = w; };
      fft_sched_lds<N, RPW, 64, -1, false, LY, SCH>(xch, twc, ld2, keepZ);
      constexpr int JOUT = (HP + NB - 1) / NB;   // slots holding a stored column k < HP
#pragma unroll
      for (int j = 0; j < JOUT; j++) {
        const double2 w = Z[R - 1 - j];
        double2 Zc = make_double2(bperm(w.x), bperm(w.y));
        if (bbL == 0) Zc = Z[(R - j) % R];
        const int k = bbL + j * NB;
        if (!act || k >= HP) continue;
        const double2 Zk = Z[j];
        double2 o0 = make_double2(0.0, 0.0), o1 = o0;
        if (k < hx) {
          if (l == 0) {
            const double zr = Zk.x, zi = Zk.y, cr = Zc.x, ci = -Zc.y;
            o0 = make_double2(0.5 * (zr + cr), 0.5 * (zi + ci));
            o1 = make_double2(0.5 * (zi - ci), -0.5 * (zr - cr));
          } else {
            o0 = make_double2(0.5 * (Zk.x + Zc.x), 0.5 * (Zk.y - Zc.y));
          }
        }
        const int m = (row0 + tL) * HP + k;
        if (l == 0) {
          P[m] = o0;
          P[N * HP + m] = o1;
        } else {
          P[2 * N * HP + m] = o0;
        }
      }
    }
  }
  // maxima of the plane into the control block (xch of the first wave as the reduction's scratch)
  block_atomic_max2<NT>(om2, u2, &ctl->om2, &ctl->u2, reinterpret_cast<double *>(lds_raw + C::PLANE));
  // (block_atomic_max2 ends with the workgroup past a barrier: every row's products are in P)
  __syncthreads();
  // 3. y-forward of the three products: P[c][y][kx] -> X4[c][z][cy][kx] = X1[c][z][cy][kx], retained ky only
  if constexpr (!C::PER_FIELD) {
    using LY3 = PlaneLY<N, HP, 3>;
    auto load = [&](int n, int t) -> double2 { return P[LY3::at(n, t)]; };
    auto store = [&](int n, int t, double2 v) {
      const int cy = compact_of(n, g);
      if (cy >= 0) X1[((size_t)(t / HP) * N + z) * plane + (size_t)cy * g.hp + t % HP] = v;
    };
    fft<N, LY3::T, NT, -1, true, LY3>(P, tw, load, store);
  } else {   // rows of the hx retained columns: the padding columns of X4 keep the zeros the z-inverse wrote there
    using LY = PlaneLY<N, HP, 1>;
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
      double2 *Pc = P + c * N * HP;
      auto load = [&](int n, int t) -> double2 { return Pc[LY::at(n, t)]; };
      auto store = [&](int n, int t, double2 v) {
        const int cy = compact_of(n, g);
        if (cy >= 0) X1[((size_t)c * N + z) * plane + (size_t)cy * g.hp + t] = v;
      };
      fft<N, LY::T, NT, -1, true, LY>(Pc, tw, load, store);
    }
  }
}

}  // namespace fast
}  // namespace nsl
