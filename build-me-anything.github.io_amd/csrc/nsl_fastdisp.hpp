// Host dispatch of the size-specialised pass kernels (nsl_fastk.hpp). Included by nslab_hip.hip.
// Sizes with a compiled radix schedule run the fast kernels; every other valid N runs the generic engine.
#pragma once
#include "nsl_fastk.hpp"

// Lines per workgroup in the pencil passes (T) and row pairs per workgroup in the x pass (LP), per N.
#ifndef NSL_T_SMALL
#define NSL_T_SMALL 8
#endif
#ifndef NSL_ZINV_UW_MAXN
#define NSL_ZINV_UW_MAXN 1024
#endif
#ifndef NSL_LP_ELEMS
#define NSL_LP_ELEMS 4096
#endif
// Up to 64^3 one tile covers the whole row pitch (16, 24, 24 lines), so a workgroup's 256 threads have work: in a
// pool of 32 solves the 48^3 step went from 1.29 to 0.98 ms, 32^3 from 0.56 to 0.53; at 96^3 (40 lines) it was 5 %
// slower. 128^3 takes 16 lines (+3.5 % alone, +3.9 % in a pool); 24 lines at 192^3 were 4.6 % slower alone.
static constexpr int fast_T(int N) {
  return N <= 64 ? ((N / 3 + 1) + 7) & ~7 : N == 128 ? 16 : N <= 256 ? NSL_T_SMALL : 8;
}
// The other tile width the autotuner may pick (pt = 1), or 0: it must divide the row pitch hp (whole tiles) and fit
// LDS (N x T complex). (16 lines at 512^3 would spill the z-inverse: 476 bytes of scratch per lane.)
static constexpr int alt_T(int N) {
  return N == 32 || N == 48 || N == 64 ? 8 : N == 96 ? 40 : N == 128 ? 24 : N == 192 ? 24 : 0;
}
// Compute units of the current device (cached per device).
static int device_cus() {
  static std::mutex mu;
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  std::lock_guard<std::mutex> lk(mu);
  if (cache[dev] == 0) {
    int n = 0;
    cache[dev] = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0 ? n : 256;
  }
  return cache[dev];
}
static constexpr int fast_LP(int N) {
  return 6 * N * 8 <= NSL_LP_ELEMS ? 8 : 6 * N * 4 <= NSL_LP_ELEMS ? 4 : 6 * N * 2 <= NSL_LP_ELEMS ? 2 : 1;
}

#define NSL_FAST_SIZES(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

static TuneCfg tune_for(int N);
static bool fast_setup(Xform &x) {
  switch (x.g.N) {
#define NSL_CASE(n) case n:
    NSL_FAST_SIZES(NSL_CASE)
#undef NSL_CASE
    x.fast = true;
    x.tune = tune_for(x.g.N);
    x.fT = x.tune.pt ? alt_T(x.g.N) : fast_T(x.g.N);
    return true;
    default:
      x.fast = false;
      return false;
  }
}

static int fast_ntile(const Xform &x) { return (x.g.hx + x.fT - 1) / x.fT; }

template <int N, int T, int MODE>
static void kf_zinv_launch(const Xform &x, const double2 *U, double2 *X1, int nout, int gc, const Ctl *ctl, const Batch &bt) {
  const int lds = N * T * (int)sizeof(double2);
  allow_lds(fast::kf_zinv<N, T, MODE>, lds);
  hipLaunchKernelGGL((fast::kf_zinv<N, T, MODE>), dim3((unsigned)(nout * fast_ntile(x) * x.g.by), 1, bt.nz), dim3(fast::PassNT<N>::v), lds, x.st, U, X1, x.g, x.twN, nout, gc, ctl, bt.sb);
}

template <int N, int T>
static int fast_zinv_nt(const Xform &x, const double2 *U, double2 *X1, int mode, int nout, int gc, const Ctl *ctl, const Batch &bt,
                        const Fin &fin) {
  const int lds = N * T * (int)sizeof(double2);
  if (mode == 1 && nout == 6 && N <= NSL_ZINV_UW_MAXN) {   // the rhs: u and omega from five transforms per tile
    // (4-line tiles at 1024^3, two workgroups per CU: 12.1 against 8.8 ms)
    if (fin.nfin > 0 && (ctl == nullptr || fast::ZinvFin<N, T>::CH == 0))
      return fail(NSL_ERR_STATE, "no rhs booking in this z-inverse");
    allow_lds(fast::kf_zinv_uw<N, T>, lds);
    // the booking workgroup writes the control block: a launch with fin.nfin > 0 comes from a solver's own rhs
    hipLaunchKernelGGL((fast::kf_zinv_uw<N, T>), dim3((unsigned)(fast_ntile(x) * 2 * x.g.by + (fin.nfin > 0 ? 1 : 0)), 1, bt.nz),
                       dim3(fast::ZinvNT<N>::v), lds, x.st, U, X1, x.g, x.twN, const_cast<Ctl *>(ctl), bt.sb, fin.partials,
                       fin.nfin, fin.fmode);
    LAUNCH_CHECK();
    return NSL_OK;
  }
  if (fin.nfin > 0) return fail(NSL_ERR_STATE, "no rhs booking in this z-inverse");
  if (mode == 0) kf_zinv_launch<N, T, 0>(x, U, X1, nout, gc, ctl, bt);
  else if (mode == 1) kf_zinv_launch<N, T, 1>(x, U, X1, nout, gc, ctl, bt);
  else kf_zinv_launch<N, T, 2>(x, U, X1, nout, gc, ctl, bt);
  LAUNCH_CHECK();
  return NSL_OK;
}
template <int N, int T>
static int fast_yinv_nt(const Xform &x, const double2 *X1, double2 *X2, int ncomp, const Ctl *ctl, const Batch &bt) {
  const int lds = N * T * (int)sizeof(double2);
  allow_lds(fast::kf_yinv<N, T>, lds);
  hipLaunchKernelGGL((fast::kf_yinv<N, T>), dim3((unsigned)(fast_ntile(x) * N), (unsigned)ncomp, bt.nz), dim3(fast::PassNT<N>::v), lds, x.st, X1, X2, x.g, x.twN, ctl, bt.sb);
  LAUNCH_CHECK();
  return NSL_OK;
}
template <int N, int T>
static int fast_yfwd_nt(const Xform &x, const double2 *X3, double2 *X4, int ncomp, const Ctl *ctl, const Batch &bt,
                       const double *pmax, int nblk) {
  const int lds = N * T * (int)sizeof(double2);
  allow_lds(fast::kf_yfwd<N, T>, lds);
  const unsigned grid = (unsigned)(fast_ntile(x) * N);
  // workgroups that combine the x-pass maxima: about 8 KB of them each, at most the grid's first row
  const int nred = pmax && nblk > 0 ? (int)std::min<long long>(grid, nblk / fast::PassNT<N>::v + 1) : 0;
  hipLaunchKernelGGL((fast::kf_yfwd<N, T>), dim3(grid, (unsigned)ncomp, bt.nz), dim3(fast::PassNT<N>::v), lds, x.st, X3, X4, x.g, x.twN, ctl, bt.sb, pmax, nblk, nred);
  LAUNCH_CHECK();
  return NSL_OK;
}
template <int N, int T>
static int fast_zfwd_nt(const Xform &x, const double2 *X4, double2 *OUT, int ncomp, const Ctl *ctl, const Batch &bt) {
  const int lds = N * T * (int)sizeof(double2);
  const double inv = 1.0 / ((double)N * N * N);
  allow_lds(fast::kf_zfwd<N, T>, lds);
  hipLaunchKernelGGL((fast::kf_zfwd<N, T>), dim3((unsigned)(fast_ntile(x) * x.g.by), (unsigned)ncomp, bt.nz), dim3(fast::PassNT<N>::v), lds, x.st, X4, OUT, x.g, x.twN, inv, ctl, bt.sb);
  LAUNCH_CHECK();
  return NSL_OK;
}
// The pencil passes at the solver's tile width x.fT (fast_T(N), or alt_T(N) when the autotuner picked it).
template <int N>
static int fast_zinv_n(const Xform &x, const double2 *U, double2 *X1, int mode, int nout, int gc, const Ctl *ctl, const Batch &bt,
                       const Fin &fin) {
  if constexpr (alt_T(N) > 0) if (x.fT == alt_T(N)) return fast_zinv_nt<N, alt_T(N)>(x, U, X1, mode, nout, gc, ctl, bt, fin);
  return fast_zinv_nt<N, fast_T(N)>(x, U, X1, mode, nout, gc, ctl, bt, fin);
}
template <int N>
static int fast_yinv_n(const Xform &x, const double2 *X1, double2 *X2, int ncomp, const Ctl *ctl, const Batch &bt) {
  if constexpr (alt_T(N) > 0) if (x.fT == alt_T(N)) return fast_yinv_nt<N, alt_T(N)>(x, X1, X2, ncomp, ctl, bt);
  return fast_yinv_nt<N, fast_T(N)>(x, X1, X2, ncomp, ctl, bt);
}
template <int N>
static int fast_yfwd_n(const Xform &x, const double2 *X3, double2 *X4, int ncomp, const Ctl *ctl, const Batch &bt,
                       const double *pmax, int nblk) {
  if constexpr (alt_T(N) > 0) if (x.fT == alt_T(N)) return fast_yfwd_nt<N, alt_T(N)>(x, X3, X4, ncomp, ctl, bt, pmax, nblk);
  return fast_yfwd_nt<N, fast_T(N)>(x, X3, X4, ncomp, ctl, bt, pmax, nblk);
}
template <int N>
static int fast_zfwd_n(const Xform &x, const double2 *X4, double2 *OUT, int ncomp, const Ctl *ctl, const Batch &bt) {
  if constexpr (alt_T(N) > 0) if (x.fT == alt_T(N)) return fast_zfwd_nt<N, alt_T(N)>(x, X4, OUT, ncomp, ctl, bt);
  return fast_zfwd_nt<N, fast_T(N)>(x, X4, OUT, ncomp, ctl, bt);
}

// x-pass forms compiled for each N. Row pairs: N <= 384, LP row pairs per workgroup with 6*LP*N complex of LDS
// (at most 160 KB). Rows: N >= 384. One wave per row group (plain or staged): N = 256, 512.
// NSL_XP_PLANE: the y-inverse, the x pass and the y-forward as one pass, one workgroup per z-plane (fast::kf_plane).
enum { NSL_XP_PAIR = 1, NSL_XP_ROWS = 2, NSL_XP_WAVE = 3, NSL_XP_STAGED = 4, NSL_XP_PLANE = 5 };
static constexpr bool xp_pair_lp_ok(int N, int LP) { return N <= 384 && 6 * LP * N * 16 <= 160 * 1024; }
static constexpr bool xp_has(int N, int form) {
  return form == NSL_XP_PAIR ? N <= 384
       : form == NSL_XP_ROWS ? N >= 384
       : form == NSL_XP_WAVE ? (N == 256 || N == 512)
       : form == NSL_XP_STAGED ? (N == 32 || N == 48 || N == 64 || N == 96 || N == 128 || N == 192 || N == 256 ||
                                  N == 384 || N == 512 || N == 1024)
       : form == NSL_XP_PLANE ? fast::plane_has(N)
       : false;
}
// sizes with a variant-1 schedule of the staged x pass (XnlSched) and of the fused update (ZuSchedV)
static constexpr bool xs_has(int N) {
  return N == 32 || N == 64 || N == 96 || N == 128 || N == 192 || N == 256 || N == 384 || N == 512 || N == 1024;
}
// sizes with the one-wave shape of the staged x pass (variant 2: fast::XnlOneWave)
static constexpr bool xs2_has(int N) { return N == 48 || N == 96 || N == 192 || N == 384; }
static constexpr bool zs_has(int N) {
  return N == 32 || N == 48 || N == 64 || N == 96 || N == 128 || N == 192 || N == 256 || N == 384 || N == 512 ||
         N == 1024;
}
static constexpr bool zu_has(int N) {
  return N == 32 || N == 48 || N == 64 || N == 96 || N == 128 || N == 192 || N == 256 || N == 384 || N == 512 ||
         N == 1024;
}

// Workgroups the x pass of form c launches at grid N on a device of `cus` compute units; each leaves one maxima pair
// (max |omega|^2, max |u|^2) in pmax. 0 when the form is not compiled for N.
template <int N>
static long long xnl_workgroups(const TuneCfg &c, int cus) {
  switch (c.xpass) {
    case NSL_XP_STAGED:
      if constexpr (xp_has(N, NSL_XP_STAGED)) {
        int RPW = fast::XnlStaged<N, 0>::RPW, WPR = fast::XnlStaged<N, 0>::WPR;
        if (c.xs == 1) { RPW = fast::XnlStaged<N, 1>::RPW; WPR = fast::XnlStaged<N, 1>::WPR; }
        if constexpr (xs2_has(N)) if (c.xs == 2) { RPW = fast::XnlStaged<N, 2>::RPW; WPR = fast::XnlStaged<N, 2>::WPR; }
        long long grid = (long long)N * N / RPW;
        // persistent: xwaves waves per CU walk the row groups
        if (c.xwaves > 0) grid = std::min<long long>(grid, std::max(1, cus * c.xwaves / WPR));
        return grid;
      }
      return 0;
    case NSL_XP_WAVE:
      if constexpr (xp_has(N, NSL_XP_WAVE)) return (long long)N * N / (64 / (N / 8));
      return 0;
    case NSL_XP_PLANE:   // one workgroup per z-plane; its maxima go to the control block (no pairs in pmax)
      if constexpr (xp_has(N, NSL_XP_PLANE)) return N;
      return 0;
    case NSL_XP_ROWS:
      if constexpr (xp_has(N, NSL_XP_ROWS)) return (long long)N * N;   // one row per workgroup
      return 0;
    case NSL_XP_PAIR:
      if (c.lp == 1 || c.lp == 2 || c.lp == 4 || c.lp == 8) {
        const bool ok = c.lp == 1 ? xp_pair_lp_ok(N, 1) : c.lp == 2 ? xp_pair_lp_ok(N, 2) : c.lp == 4 ? xp_pair_lp_ok(N, 4) : xp_pair_lp_ok(N, 8);
        if (ok) return ((long long)N * (N / 2) + c.lp - 1) / c.lp;
      }
      return 0;
    default:
      return 0;
  }
}

// The x pass's launch plan, checked before anything is enqueued: the form must be compiled for N and its
// workgroups' maxima pairs must fit the cap_pairs the slot's pmax buffer holds (N^2 pairs).
template <int N>
static int xnl_plan(const TuneCfg &c, int cus, long long cap_pairs, long long *grid) {
  const long long g = xnl_workgroups<N>(c, cus);
  if (g <= 0) return fail(NSL_ERR_STATE, "x-pass form " + std::to_string(c.xpass) + "/" + std::to_string(c.lp) + " is not compiled for N = " + std::to_string(N));
  if (g > cap_pairs) return fail(NSL_ERR_STATE, "x-pass maxima exceed their buffer (" + std::to_string(g) + " workgroups, " + std::to_string(cap_pairs) + " pairs)");
  *grid = g;
  return NSL_OK;
}

template <int N, int LP>
static void fast_xnl_pair_launch(const Xform &x, double2 *X2, double *pmax, Ctl *ctl, const Batch &bt, long long grid) {
  if constexpr (xp_pair_lp_ok(N, LP)) {
    const int lds = fast::XnlLayout<N, 6 * LP>::SIZE * (int)sizeof(double2);
    allow_lds(fast::kf_xnl<N, LP>, lds);
    hipLaunchKernelGGL((fast::kf_xnl<N, LP>), dim3((unsigned)grid, 1, bt.nz), dim3(NSL_NT), lds, x.st, X2, x.g, x.twN, pmax, ctl, bt.sb);
  }
}

// The x pass leaves one maxima pair per workgroup in pmax (N^2 pairs per slot); *nred_out is their number, which the
// y-forward pass combines (fast_yfwd_n). The plan is checked before the launch.
template <int N>
static int fast_xnl_n(const Xform &x, double2 *X2, double *pmax, Ctl *ctl, const Batch &bt, int *nred_out) {
  const TuneCfg &c = x.tune;
  long long grid = 0;
  TRY(xnl_plan<N>(c, device_cus(), (long long)N * N, &grid));
  switch (c.xpass) {
    case NSL_XP_STAGED:
      if constexpr (xp_has(N, NSL_XP_STAGED)) {
        if (c.xs == 0) {
          hipLaunchKernelGGL((fast::kf_xnl_dma<N, 0>), dim3((unsigned)grid, 1, bt.nz), dim3(fast::XnlStaged<N, 0>::NT), 0, x.st, X2, x.g, x.twN, pmax, ctl, bt.sb);
        } else if (c.xs == 1) {
          if constexpr (xs_has(N))
            hipLaunchKernelGGL((fast::kf_xnl_dma<N, 1>), dim3((unsigned)grid, 1, bt.nz), dim3(fast::XnlStaged<N, 1>::NT), 0, x.st, X2, x.g, x.twN, pmax, ctl, bt.sb);
        } else {
          if constexpr (xs2_has(N))
            hipLaunchKernelGGL((fast::kf_xnl_dma<N, 2>), dim3((unsigned)grid, 1, bt.nz), dim3(fast::XnlStaged<N, 2>::NT), 0, x.st, X2, x.g, x.twN, pmax, ctl, bt.sb);
        }
      }
      break;
    case NSL_XP_WAVE:
      if constexpr (xp_has(N, NSL_XP_WAVE)) {
        constexpr int RPW = 64 / (N / 8);
        const int lds = fast::LineMajor<N, RPW>::SIZE * (int)sizeof(double2);
        hipLaunchKernelGGL((fast::kf_xnl_wave<N>), dim3((unsigned)grid, 1, bt.nz), dim3(64), lds, x.st, X2, x.g, x.twN, pmax, ctl, bt.sb);
      }
      break;
    case NSL_XP_ROWS:
      if constexpr (xp_has(N, NSL_XP_ROWS)) {
        constexpr int RP = 1, NT = 192;   // 3 lines of N/8 or N/16 first-stage butterflies: one per thread
        const int lds = N * 3 * RP * (int)sizeof(double2);
        allow_lds(fast::kf_xnl_rows<N, RP, NT>, lds);
        hipLaunchKernelGGL((fast::kf_xnl_rows<N, RP, NT>), dim3((unsigned)grid, 1, bt.nz), dim3(NT), lds, x.st, X2, x.g, x.twN, pmax, ctl, bt.sb);
      }
      break;
    default:   // NSL_XP_PAIR (the plan accepted lp)
      switch (c.lp) {
        case 1: fast_xnl_pair_launch<N, 1>(x, X2, pmax, ctl, bt, grid); break;
        case 2: fast_xnl_pair_launch<N, 2>(x, X2, pmax, ctl, bt, grid); break;
        case 4: fast_xnl_pair_launch<N, 4>(x, X2, pmax, ctl, bt, grid); break;
        default: fast_xnl_pair_launch<N, 8>(x, X2, pmax, ctl, bt, grid); break;
      }
      break;
  }
  LAUNCH_CHECK();
  *nred_out = (int)grid;   // the maxima pairs the y-forward pass combines (kf_yfwd)
  return NSL_OK;
}

#define NSL_DISPATCH(fn, ...)                                          \
  switch (x.g.N) {                                                     \
    case 16: return fn<16>(x, __VA_ARGS__);                            \
    case 24: return fn<24>(x, __VA_ARGS__);                            \
    case 32: return fn<32>(x, __VA_ARGS__);                            \
    case 48: return fn<48>(x, __VA_ARGS__);                            \
    case 64: return fn<64>(x, __VA_ARGS__);                            \
    case 96: return fn<96>(x, __VA_ARGS__);                            \
    case 128: return fn<128>(x, __VA_ARGS__);                          \
    case 192: return fn<192>(x, __VA_ARGS__);                          \
    case 256: return fn<256>(x, __VA_ARGS__);                          \
    case 384: return fn<384>(x, __VA_ARGS__);                          \
    case 512: return fn<512>(x, __VA_ARGS__);                          \
    case 768: return fn<768>(x, __VA_ARGS__);                          \
    case 1024: return fn<1024>(x, __VA_ARGS__);                        \
    default: return fail(NSL_ERR_STATE, "no size-specialised kernels for this N"); \
  }

static int fast_dispatch_zinv(const Xform &x, const double2 *U, double2 *X1, int mode, int nout, int gc, const Ctl *ctl, const Batch &bt,
                              const Fin &fin) {
  NSL_DISPATCH(fast_zinv_n, U, X1, mode, nout, gc, ctl, bt, fin)
}
// P2-P4 in one pass (the x-pass form NSL_XP_PLANE): X1's six fields of every z-plane in, X4 (X1's first three
// fields) out, the maxima into the control block.
template <int N>
static int fast_plane_n(const Xform &x, double2 *X1, Ctl *ctl, const Batch &bt) {
  if constexpr (fast::plane_has(N)) {
    constexpr int lds = fast::PlaneCfg<N>::LDS_BYTES;
    allow_lds(fast::kf_plane<N>, lds);
    hipLaunchKernelGGL((fast::kf_plane<N>), dim3((unsigned)N, 1, bt.nz), dim3(fast::PlaneCfg<N>::NT), lds, x.st, X1, x.g,
                       x.twN, ctl, bt.sb);
    LAUNCH_CHECK();
    return NSL_OK;
  }
  return fail(NSL_ERR_STATE, "the plane pass is not compiled for N = " + std::to_string(N));
}
static bool fast_has_plane(const Xform &x) { return x.fast && x.tune.xpass == NSL_XP_PLANE; }

static int fast_dispatch_yinv(const Xform &x, const double2 *X1, double2 *X2, int ncomp, const Ctl *ctl, const Batch &bt) {
  NSL_DISPATCH(fast_yinv_n, X1, X2, ncomp, ctl, bt)
}
static int fast_dispatch_yfwd(const Xform &x, const double2 *X3, double2 *X4, int ncomp, const Ctl *ctl, const Batch &bt,
                              const double *pmax, int nblk) {
  NSL_DISPATCH(fast_yfwd_n, X3, X4, ncomp, ctl, bt, pmax, nblk)
}
static int fast_dispatch_plane(const Xform &x, double2 *X1, Ctl *ctl, const Batch &bt) {
  NSL_DISPATCH(fast_plane_n, X1, ctl, bt)
}
static int fast_dispatch_zfwd(const Xform &x, const double2 *X4, double2 *OUT, int ncomp, const Ctl *ctl, const Batch &bt) {
  NSL_DISPATCH(fast_zfwd_n, X4, OUT, ncomp, ctl, bt)
}
static int fast_dispatch_xnl(const Xform &x, double2 *X2, double *pmax, Ctl *ctl, const Batch &bt, int *nred_out) {
  NSL_DISPATCH(fast_xnl_n, X2, pmax, ctl, bt, nred_out)
}
static int xnl_plan_n(int N, const TuneCfg &c, int cus, long long cap_pairs, long long *grid) {
  switch (N) {
#define NSL_CASE(n) case n: return xnl_plan<n>(c, cus, cap_pairs, grid);
    NSL_FAST_SIZES(NSL_CASE)
#undef NSL_CASE
    default: return fail(NSL_ERR_ARG, "no size-specialised kernels for N = " + std::to_string(N));
  }
}

// z-forward fused with the stage-`stage` update (kf_zfwd_upd) where it is compiled (zu_has) and chosen
// (tune.zfused).
static bool fast_has_zu(const Xform &x) { return x.fast && zu_has(x.g.N) && x.tune.zfused; }
// Workgroups (and sum partials) of the fused pass.
static int fast_zu_blocks(const Geo &g) {
  const int T = fast::zu_lines(g.N);   // fast::ZuCfg<N>::T
  return ((g.hx + T - 1) / T) * g.by;
}
template <int N, int STAGE, int V>
static void fast_zu_launch_v(const Xform &x, const double2 *X4, const double2 *U, double2 *S, double2 *ACC, double2 *TS,
                             double *partials, Ctl *ctl, const Batch &bt) {
  using C = fast::ZuCfg<N>;
  const int lds = N * C::T * (int)sizeof(double2);
  const double inv = 1.0 / ((double)N * N * N);
  allow_lds(fast::kf_zfwd_upd<N, STAGE, V>, lds);
  hipLaunchKernelGGL((fast::kf_zfwd_upd<N, STAGE, V>), dim3((unsigned)fast_zu_blocks(x.g), 1, bt.nz), dim3(C::NT), lds, x.st,
                     X4, U, S, ACC, TS, x.g, x.twN, inv, partials, ctl, bt.sb);
}
// the fused update with the tuned schedule variant (zs)
template <int N, int STAGE>
static void fast_zu_launch(const Xform &x, const double2 *X4, const double2 *U, double2 *S, double2 *ACC, double2 *TS,
                           double *partials, Ctl *ctl, const Batch &bt) {
  if constexpr (zs_has(N))
    if (x.tune.zs) return fast_zu_launch_v<N, STAGE, 1>(x, X4, U, S, ACC, TS, partials, ctl, bt);
  fast_zu_launch_v<N, STAGE, 0>(x, X4, U, S, ACC, TS, partials, ctl, bt);
}
template <int N>
static int fast_zu_n(const Xform &x, int stage, const double2 *X4, const double2 *U, double2 *S, double2 *ACC,
                     double2 *TS, double *partials, Ctl *ctl, const Batch &bt) {
  switch (stage) {
    case 1: fast_zu_launch<N, 1>(x, X4, U, S, ACC, TS, partials, ctl, bt); break;
    case 2: fast_zu_launch<N, 2>(x, X4, U, S, ACC, TS, partials, ctl, bt); break;
    case 3: fast_zu_launch<N, 3>(x, X4, U, S, ACC, TS, partials, ctl, bt); break;
    default: fast_zu_launch<N, 4>(x, X4, U, S, ACC, TS, partials, ctl, bt); break;
  }
  LAUNCH_CHECK();
  return NSL_OK;
}
static int fast_dispatch_zu(const Xform &x, int stage, const double2 *X4, const double2 *U, double2 *S, double2 *ACC,
                            double2 *TS, double *partials, Ctl *ctl, const Batch &bt) {
  switch (x.g.N) {
    case 32: return fast_zu_n<32>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 48: return fast_zu_n<48>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 64: return fast_zu_n<64>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 96: return fast_zu_n<96>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 128: return fast_zu_n<128>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 192: return fast_zu_n<192>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 256: return fast_zu_n<256>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 384: return fast_zu_n<384>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 512: return fast_zu_n<512>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    case 1024: return fast_zu_n<1024>(x, stage, X4, U, S, ACC, TS, partials, ctl, bt);
    default: return fail(NSL_ERR_STATE, "no fused z-forward for this N");
  }
}
