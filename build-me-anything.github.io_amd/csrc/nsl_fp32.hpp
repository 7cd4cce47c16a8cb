// Passes of an FP32 solve: the spectral arrays in complex64 (float2) in HBM, the transforms, the curl and the product
// u x omega in single precision (float2 registers and LDS, float twiddles x.twN32), the update (projection, viscous
// term, RK4) and every sum in FP64 registers. The same size-specialised rhs kernels (nsl_fastk.hpp) instantiated with
// storage E = float2, which is also their compute type, with the z-forward fused into the update and one x-pass
// form per N. The y passes take 16-line tiles from 96^3 to 384^3: a row of 16
// complex64 is one 128-byte line, where 8 lines read half lines (at 256^3, 8 lines fetched as many bytes as FP64:
// y-forward 137 us, 16 lines 91 us; y-inverse 229 and 198 us). The z-inverse and the fused update keep the FP64
// tiles: with 16 lines they ran slower (256^3: 171 against 112 us, one wave per SIMD at 256 registers; the update
// 10 % slower). Included by nslab_hip.hip after nsl_tune.hpp. An FP32 solve is a single solve (never a pool slot),
// so every launch has gridDim.z = 1 and sb = 0.
//
// The reference's FP32 mode (nslab_gpu.py:56-65) holds every array in complex64/float32 and transforms in single
// precision. Here too: every array that lives in HBM between passes is complex64, and the transforms and the
// product run in single precision, so the state, the RK4 registers and each pass's output carry FP32 rounding as
// there. The update and the sums stay FP64 (a little less roundoff than an all-single pipeline, at no cost).
#pragma once

// Sizes with an FP32 path: those with the fused z-forward and update.
static constexpr bool fp32_has(int N) { return zu_has(N); }
// x-pass form of FP32 solves: one wave per row group at 256^3 and 512^3, one row per workgroup above 384^3, row
// pairs (fast_LP(N) per workgroup) otherwise.
static constexpr int xp32_form(int N) {
  return (N == 256 || N == 512) ? NSL_XP_WAVE : N > 384 ? NSL_XP_ROWS : NSL_XP_PAIR;
}
// Workgroups of the FP32 x pass (one maxima pair each, at most N^2).
static constexpr long long xnl32_workgroups(int N) {
  return xp32_form(N) == NSL_XP_WAVE ? (long long)N * N / (64 / (N / 8))
       : xp32_form(N) == NSL_XP_ROWS ? (long long)N * N
       : ((long long)N * (N / 2) + fast_LP(N) - 1) / fast_LP(N);
}

// Tile of the FP32 y passes: the whole row pitch up to 64^3 (as FP64), 16 lines to 384^3, 8 above (16 lines at
// N >= 512 need 128+ KB of LDS per workgroup).
static constexpr int fast_T32(int N) { return N <= 64 ? fast_T(N) : N <= 384 ? 16 : 8; }

template <int N> struct P32 {
  static constexpr int T = fast_T32(N), TZ = fast_T(N);   // y passes; z-inverse
  static constexpr Geo g = fast::trunc_geo<N>();
  static constexpr int ntile = (g.hx + T - 1) / T, ntile_z = (g.hx + TZ - 1) / TZ;

  // P1: u and omega of U along z (six fields into X1)
  // fin: the previous rhs's booking, done by one extra workgroup (kf_zinv_uw)
  static int zinv(const Xform &x, const float2 *U, float2 *X1, Ctl *ctl, const Fin &fin) {
    const int lds = N * TZ * (int)sizeof(double2);
    static_assert(fast::ZinvFin<N, TZ>::CH > 0, "the z-inverse's LDS holds the booking's scratch");
    allow_lds(fast::kf_zinv_uw<N, TZ, float2>, lds);
    hipLaunchKernelGGL((fast::kf_zinv_uw<N, TZ, float2>), dim3((unsigned)(ntile_z * 2 * g.by + (fin.nfin > 0 ? 1 : 0))),
                       dim3(fast::ZinvNT<N>::v), lds, x.st, U, X1, x.g, (const float2 *)x.twN32, ctl, (size_t)0,
                       fin.partials, fin.nfin, fin.fmode);
    LAUNCH_CHECK();
    return NSL_OK;
  }
  // P2: six fields along y
  static int yinv(const Xform &x, const float2 *X1, float2 *X2, const Ctl *ctl) {
    const int lds = N * T * (int)sizeof(float2);
    allow_lds(fast::kf_yinv<N, T, float2>, lds);
    hipLaunchKernelGGL((fast::kf_yinv<N, T, float2>), dim3((unsigned)(ntile * N), 6u), dim3(fast::PassNT<N>::v), lds, x.st,
                       X1, X2, x.g, (const float2 *)x.twN32, ctl, (size_t)0);
    LAUNCH_CHECK();
    return NSL_OK;
  }
  // P3: x C2R, u x omega, x R2C in place on X2; one maxima pair per workgroup into pmax (N^2 pairs)
  static int xnl(const Xform &x, float2 *X2, double *pmax, Ctl *ctl, int *nred) {
    constexpr long long grid = xnl32_workgroups(N);
    static_assert(grid > 0 && grid <= (long long)N * N, "x-pass maxima fit their buffer");
    if constexpr (xp32_form(N) == NSL_XP_WAVE) {
      constexpr int RPW = 64 / (N / 8);
      const int lds = fast::LineMajor<N, RPW>::SIZE * (int)sizeof(float2);
      hipLaunchKernelGGL((fast::kf_xnl_wave<N, float2>), dim3((unsigned)grid), dim3(64), lds, x.st, X2, x.g,
                         (const float2 *)x.twN32, pmax, ctl, (size_t)0);
    } else if constexpr (xp32_form(N) == NSL_XP_ROWS) {
      constexpr int RP = 1, NT = 192;
      const int lds = N * 3 * RP * (int)sizeof(float2);
      allow_lds(fast::kf_xnl_rows<N, RP, NT, float2>, lds);
      hipLaunchKernelGGL((fast::kf_xnl_rows<N, RP, NT, float2>), dim3((unsigned)grid), dim3(NT), lds, x.st, X2, x.g,
                         (const float2 *)x.twN32, pmax, ctl, (size_t)0);
    } else {
      constexpr int LP = fast_LP(N);
      static_assert(xp_pair_lp_ok(N, LP), "row pairs fit LDS");
      const int lds = fast::XnlLayout<N, 6 * LP>::SIZE * (int)sizeof(float2);
      allow_lds(fast::kf_xnl<N, LP, float2>, lds);
      hipLaunchKernelGGL((fast::kf_xnl<N, LP, float2>), dim3((unsigned)grid), dim3(NSL_NT), lds, x.st, X2, x.g,
                         (const float2 *)x.twN32, pmax, ctl, (size_t)0);
    }
    LAUNCH_CHECK();
    *nred = (int)grid;
    return NSL_OK;
  }
  // P4: the three nonlinear components along y, truncated to the retained ky; combines the x pass's maxima
  static int yfwd(const Xform &x, const float2 *X2, float2 *X1, const Ctl *ctl, const double *pmax, int nblk) {
    const int lds = N * T * (int)sizeof(float2);
    allow_lds(fast::kf_yfwd<N, T, float2>, lds);
    const unsigned grid = (unsigned)(ntile * N);
    const int nred = pmax && nblk > 0 ? (int)std::min<long long>(grid, nblk / fast::PassNT<N>::v + 1) : 0;
    hipLaunchKernelGGL((fast::kf_yfwd<N, T, float2>), dim3(grid, 3u), dim3(fast::PassNT<N>::v), lds, x.st, X2, X1, x.g,
                       (const float2 *)x.twN32, ctl, (size_t)0, pmax, nblk, nred);
    LAUNCH_CHECK();
    return NSL_OK;
  }
  // P5: z-forward fused with the stage-`stage` update
  // workgroups of the fused update (one sum partial each)
  using ZC = fast::ZuCfg<N>;
  static constexpr int zu_blocks = (g.hx + ZC::T - 1) / ZC::T * g.by;
  template <int STAGE>
  static void zu_stage(const Xform &x, const float2 *X1, const float2 *U, float2 *S, float2 *ACC, float2 *TS,
                       double *partials, Ctl *ctl) {
    // LDS: the transform's N T float2, which also holds the block sums' 8 NT doubles (N T = 8 NT)
    static_assert(N * ZC::T * sizeof(float2) >= 8 * ZC::NT * sizeof(double), "the sums' scratch fits");
    const int lds = N * ZC::T * (int)sizeof(float2);
    const double inv = 1.0 / ((double)N * N * N);
    allow_lds(fast::kf_zfwd_upd<N, STAGE, 0, float2>, lds);
    hipLaunchKernelGGL((fast::kf_zfwd_upd<N, STAGE, 0, float2>), dim3((unsigned)zu_blocks), dim3(ZC::NT), lds,
                       x.st, X1, U, S, ACC, TS, x.g, (const float2 *)x.twN32, inv, partials, ctl, (size_t)0);
  }
  // *nsum: the sum partials k_finalize adds
  static int zu(const Xform &x, int stage, const float2 *X1, const float2 *U, float2 *S, float2 *ACC, float2 *TS,
                double *partials, Ctl *ctl, int *nsum) {
    static_assert(zu_blocks <= g.by * ((g.hx + 3) / 4), "sum partials fit their buffer");
    *nsum = zu_blocks;
    switch (stage) {
      case 1: zu_stage<1>(x, X1, U, S, ACC, TS, partials, ctl); break;
      case 2: zu_stage<2>(x, X1, U, S, ACC, TS, partials, ctl); break;
      case 3: zu_stage<3>(x, X1, U, S, ACC, TS, partials, ctl); break;
      default: zu_stage<4>(x, X1, U, S, ACC, TS, partials, ctl); break;
    }
    LAUNCH_CHECK();
    return NSL_OK;
  }
};

#define NSL_FP32_DISPATCH(call)                                                   \
  switch (x.g.N) {                                                                \
    case 32: return P32<32>::call;                                                \
    case 48: return P32<48>::call;                                                \
    case 64: return P32<64>::call;                                                \
    case 96: return P32<96>::call;                                                \
    case 128: return P32<128>::call;                                              \
    case 192: return P32<192>::call;                                              \
    case 256: return P32<256>::call;                                              \
    case 384: return P32<384>::call;                                              \
    case 512: return P32<512>::call;                                              \
    case 1024: return P32<1024>::call;                                            \
    default: return fail(NSL_ERR_STATE, "no FP32 kernels for N = " + std::to_string(x.g.N)); \
  }

static int zinv32(const Xform &x, const float2 *U, float2 *X1, Ctl *ctl, const Fin &fin) { NSL_FP32_DISPATCH(zinv(x, U, X1, ctl, fin)) }
static int yinv32(const Xform &x, const float2 *X1, float2 *X2, const Ctl *ctl) { NSL_FP32_DISPATCH(yinv(x, X1, X2, ctl)) }
static int xnl32(const Xform &x, float2 *X2, double *pmax, Ctl *ctl, int *nred) { NSL_FP32_DISPATCH(xnl(x, X2, pmax, ctl, nred)) }
static int yfwd32(const Xform &x, const float2 *X2, float2 *X1, const Ctl *ctl, const double *pmax, int nblk) {
  NSL_FP32_DISPATCH(yfwd(x, X2, X1, ctl, pmax, nblk))
}
static int zu32(const Xform &x, int stage, const float2 *X1, const float2 *U, float2 *S, float2 *ACC, float2 *TS,
                double *partials, Ctl *ctl, int *nsum) {
  NSL_FP32_DISPATCH(zu(x, stage, X1, U, S, ACC, TS, partials, ctl, nsum))
}
