// libnslab_hip.so — host side of the NSLab gfx950 solver and its C ABI (include/nslab_hip.h).
//
// Device memory per solver (h = floor(N/3)+1 retained kx, b = 2 floor(N/3)+1 retained ky and kz, D = b*b*h):
//   S, T, ACC, NL   3*D complex each      state, RK4 stage state, RK4 accumulator, projected nonlinear term
//   X1              6*N*b*h complex        z-transformed u and omega; later the y-forward output
//   X2              6*N*N*h complex        y-transformed u and omega; P3 overwrites its first 3 with u x omega
// All complex FP64, interleaved (double2). Modes outside the 2/3 mask are never stored: every path of the
// reference masks them to exact zero (nslab.js:235, 267, 272, 284, 289).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nslab_hip.h"
#include "nsl_kernels.hpp"

using namespace nsl;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return fail(NSL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH_CHECK()                                                                    \
  do {                                                                                    \
    hipError_t e_ = hipGetLastError();                                                    \
    if (e_ != hipSuccess) return fail(NSL_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

#define TRY(expr)                \
  do {                           \
    int r_ = (expr);             \
    if (r_ != NSL_OK) return r_; \
  } while (0)

static const double TWO_PI = 6.283185307179586;

// ------------------------------------------------------------------------------------------------ plans
static int factorize_host(int N, std::vector<int> &f) {
  f.clear();
  int n = N;
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  while (n % 2 == 0) { f.push_back(2); n /= 2; }
  while (n % 3 == 0) { f.push_back(3); n /= 3; }
  if (n != 1 || N < 4) return -1;
  return (int)f.size();
}

static std::string bad_n_msg(int N) { return "grid size must be 2^a·3^b ≥ 4 (got " + std::to_string(N) + ")"; }

struct HostPlan {
  DPlan P;
  double2 *twd = nullptr;
};

// Twiddles as fftPlan builds them (nslab.js:36-44): stage with sub-length n, m = n/r, entry (j-1)*m + p holds
// cos, sin of 2*pi*j*p/n.
static int plan_create(int N, HostPlan &hp) {
  std::vector<int> f;
  if (factorize_host(N, f) < 0) return fail(NSL_ERR_ARG, bad_n_msg(N));
  if ((int)f.size() > NSL_MAXST) return fail(NSL_ERR_ARG, "grid size too large");
  std::vector<double2> tw;
  DPlan P;
  memset(&P, 0, sizeof P);
  P.N = N;
  P.nst = (int)f.size();
  int n = N;
  for (int s = 0; s < P.nst; s++) {
    const int r = f[s], m = n / r;
    P.r[s] = r;
    P.tw[s] = (int)tw.size();
    for (int j = 1; j < r; j++)
      for (int p = 0; p < m; p++) {
        const double a = TWO_PI * j * p / n;
        tw.push_back(make_double2(cos(a), sin(a)));
      }
    n = m;
  }
  if (tw.empty()) tw.push_back(make_double2(1.0, 0.0));
  HIP_TRY(hipMalloc(&hp.twd, tw.size() * sizeof(double2)));
  HIP_TRY(hipMemcpy(hp.twd, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
  P.twd = hp.twd;
  hp.P = P;
  return NSL_OK;
}

// Tiling of the pencil passes: T lines per workgroup (power of two, T*N <= TN elements of LDS), and LP row pairs per
// workgroup in the x passes (6*LP lines of N).
struct Tiling {
  int TNp, lgT;   // pencil passes
  int TNx, LP;    // x passes
};

static bool tiling_for(int N, Tiling &t) {
  t.TNp = N <= 512 ? 4096 : 8192;
  int T = 1;
  while (T * 2 <= 16 && T * 2 * N <= t.TNp) T *= 2;
  if (T * N > t.TNp) return false;
  t.lgT = 0;
  while ((1 << t.lgT) < T) t.lgT++;
  t.TNx = 6 * N <= 4096 ? 4096 : 8192;
  if (6 * N > t.TNx) return false;
  t.LP = 1;
  while (t.LP * 2 <= 8 && 6 * N * t.LP * 2 <= t.TNx) t.LP *= 2;
  return true;
}

// Raises a kernel's dynamic LDS limit once per (kernel, size); later launches (and stream capture) skip the call.
template <typename K>
static void allow_lds(K kernel, int bytes) {
  if (bytes <= 65536) return;
  static std::mutex mu;
  static std::set<std::pair<const void *, int>> done;
  std::lock_guard<std::mutex> lk(mu);
  if (done.insert({(const void *)kernel, bytes}).second)
    (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// ------------------------------------------------------------------------------------------------ transform set
// A geometry + plan + tiling + work buffers, able to run the passes on a stream.
// Kernel forms of one N, the autotuner's choice (nsl_tune.hpp): x-pass form (NSL_XP_*), row pairs per workgroup
// of the row-pair form, persistent waves per CU of the staged form (0: one row group per workgroup), and whether
// the z-forward runs fused with the update.
struct TuneCfg {
  int xpass = 0, lp = 0, xwaves = 0, zfused = 0;
  int pt = 0, xs = 0, zs = 0;   // pencil tile width (0 default, 1 alternative); schedule variants of the staged
                                // x pass and of the fused update
};

struct Xform {
  Geo g;
  HostPlan hp;
  Tiling tl;
  hipStream_t st = 0;
  // size-specialised engine (nsl_fastk.hpp) when N has a compiled schedule
  bool fast = false;
  int fT = 0;
  TuneCfg tune;             // fixed when the transform set is made: every solve of this set runs the same kernels
  double2 *twN = nullptr;   // (cos, sin)(2 pi k / N), k < N
  float2 *twN32 = nullptr;  // the same rounded to float, for the single-precision transforms of FP32 solves
};

// Slot batching of a launch: nz slots (gridDim.z), slot buffers sb bytes apart. {1, 0}: one solve.
struct Batch {
  unsigned nz = 1;
  size_t sb = 0;
};
// The booking of the previous rhs that a fused update left pending (Ctl::pend), done by one extra workgroup of the
// next rhs's z-inverse: the update's partials, their number (0: nothing to book here) and the diagnostics stage.
struct Fin {
  const double *partials = nullptr;
  int nfin = 0;
  int fmode = 0;
};

static int fast_dispatch_zinv(const Xform &x, const double2 *U, double2 *X1, int mode, int nout, int gc, const Ctl *ctl, const Batch &bt,
                              const Fin &fin);
static int fast_dispatch_yinv(const Xform &x, const double2 *X1, double2 *X2, int ncomp, const Ctl *ctl, const Batch &bt);
static int fast_dispatch_yfwd(const Xform &x, const double2 *X3, double2 *X4, int ncomp, const Ctl *ctl, const Batch &bt,
                              const double *pmax, int nblk);
static int fast_dispatch_zfwd(const Xform &x, const double2 *X4, double2 *OUT, int ncomp, const Ctl *ctl, const Batch &bt);
static int fast_dispatch_xnl(const Xform &x, double2 *X2, double *pmax, Ctl *ctl, const Batch &bt, int *nred_out);
static bool fast_setup(Xform &x);

static Geo geo_truncated(int N) {
  Geo g;
  const int kc = N / 3;
  g.N = N; g.NH = N / 2 + 1; g.hx = kc + 1; g.by = 2 * kc + 1; g.hy = kc + 1;
  g.hp = (g.hx + 7) & ~7;   // rows start on 128-byte boundaries
  g.D = (long long)g.by * g.by * g.hp;
  return g;
}
static Geo geo_full(int N) {
  Geo g;
  g.N = N; g.NH = N / 2 + 1; g.hx = g.NH; g.by = N; g.hy = N / 2;
  g.hp = g.hx;              // the reference layout, unpadded
  g.D = (long long)g.by * g.by * g.hp;
  return g;
}

static int ntile(const Xform &x) { return (x.g.hx + (1 << x.tl.lgT) - 1) >> x.tl.lgT; }

template <int TN>
static int zinv_t(const Xform &x, const double2 *U, double2 *X1, int mode, int nout, int gc, const Ctl *ctl) {
  const int lds = (x.g.N << x.tl.lgT) * (int)sizeof(double2);
  dim3 grid((unsigned)(nout * ntile(x) * x.g.by));
  if (mode == 0) { allow_lds(k_zinv<TN, 0>, lds); hipLaunchKernelGGL((k_zinv<TN, 0>), grid, dim3(NSL_NT), lds, x.st, U, X1, x.g, x.hp.P, x.tl.lgT, nout, gc, ctl); }
  else if (mode == 1) { allow_lds(k_zinv<TN, 1>, lds); hipLaunchKernelGGL((k_zinv<TN, 1>), grid, dim3(NSL_NT), lds, x.st, U, X1, x.g, x.hp.P, x.tl.lgT, nout, gc, ctl); }
  else { allow_lds(k_zinv<TN, 2>, lds); hipLaunchKernelGGL((k_zinv<TN, 2>), grid, dim3(NSL_NT), lds, x.st, U, X1, x.g, x.hp.P, x.tl.lgT, nout, gc, ctl); }
  LAUNCH_CHECK();
  return NSL_OK;
}
// The generic engine serves single solves only (pools require a size with compiled kernels).
static int need_single(const Batch &bt) {
  return bt.nz == 1 ? NSL_OK : fail(NSL_ERR_STATE, "slot batching needs the size-specialised kernels");
}

static int zinv(const Xform &x, const double2 *U, double2 *X1, int mode, int nout, int gc, const Ctl *ctl,
                const Batch &bt = Batch(), const Fin &fin = Fin()) {
  if (x.fast) return fast_dispatch_zinv(x, U, X1, mode, nout, gc, ctl, bt, fin);
  if (fin.nfin > 0) return fail(NSL_ERR_STATE, "a pending rhs booking needs the fused update's z-inverse");
  TRY(need_single(bt));
  return x.tl.TNp == 4096 ? zinv_t<4096>(x, U, X1, mode, nout, gc, ctl) : zinv_t<8192>(x, U, X1, mode, nout, gc, ctl);
}

template <int TN>
static int yinv_t(const Xform &x, const double2 *X1, double2 *X2, int ncomp, const Ctl *ctl) {
  const int lds = (x.g.N << x.tl.lgT) * (int)sizeof(double2);
  allow_lds(k_yinv<TN>, lds);
  hipLaunchKernelGGL((k_yinv<TN>), dim3((unsigned)(ntile(x) * x.g.N), (unsigned)ncomp), dim3(NSL_NT), lds, x.st, X1, X2, x.g, x.hp.P, x.tl.lgT, ctl);
  LAUNCH_CHECK();
  return NSL_OK;
}
static int yinv(const Xform &x, const double2 *X1, double2 *X2, int ncomp, const Ctl *ctl, const Batch &bt = Batch()) {
  if (x.fast) return fast_dispatch_yinv(x, X1, X2, ncomp, ctl, bt);
  TRY(need_single(bt));
  return x.tl.TNp == 4096 ? yinv_t<4096>(x, X1, X2, ncomp, ctl) : yinv_t<8192>(x, X1, X2, ncomp, ctl);
}

template <int TN>
static int yfwd_t(const Xform &x, const double2 *X3, double2 *X4, int ncomp, const Ctl *ctl) {
  const int lds = (x.g.N << x.tl.lgT) * (int)sizeof(double2);
  allow_lds(k_yfwd<TN>, lds);
  hipLaunchKernelGGL((k_yfwd<TN>), dim3((unsigned)(ntile(x) * x.g.N), (unsigned)ncomp), dim3(NSL_NT), lds, x.st, X3, X4, x.g, x.hp.P, x.tl.lgT, ctl);
  LAUNCH_CHECK();
  return NSL_OK;
}
// pmax, nblk: x-pass maxima still to combine into ctl (the fast path's x pass leaves them; see xnl), or none
static int yfwd(const Xform &x, const double2 *X3, double2 *X4, int ncomp, const Ctl *ctl, const Batch &bt = Batch(),
                const double *pmax = nullptr, int nblk = 0) {
  if (x.fast) return fast_dispatch_yfwd(x, X3, X4, ncomp, ctl, bt, pmax, nblk);
  TRY(need_single(bt));
  return x.tl.TNp == 4096 ? yfwd_t<4096>(x, X3, X4, ncomp, ctl) : yfwd_t<8192>(x, X3, X4, ncomp, ctl);
}

template <int TN>
static int zfwd_t(const Xform &x, const double2 *X4, double2 *OUT, int ncomp, const Ctl *ctl) {
  const int lds = (x.g.N << x.tl.lgT) * (int)sizeof(double2);
  const double inv = 1.0 / ((double)x.g.N * x.g.N * x.g.N);
  allow_lds(k_zfwd<TN>, lds);
  hipLaunchKernelGGL((k_zfwd<TN>), dim3((unsigned)(ntile(x) * x.g.by), (unsigned)ncomp), dim3(NSL_NT), lds, x.st, X4, OUT, x.g, x.hp.P, x.tl.lgT, inv, ctl);
  LAUNCH_CHECK();
  return NSL_OK;
}
static int zfwd(const Xform &x, const double2 *X4, double2 *OUT, int ncomp, const Ctl *ctl, const Batch &bt = Batch()) {
  if (x.fast) return fast_dispatch_zfwd(x, X4, OUT, ncomp, ctl, bt);
  TRY(need_single(bt));
  return x.tl.TNp == 4096 ? zfwd_t<4096>(x, X4, OUT, ncomp, ctl) : zfwd_t<8192>(x, X4, OUT, ncomp, ctl);
}

static unsigned xgrid(const Xform &x) {
  const long long npairs = (long long)x.g.N * (x.g.N / 2);
  return (unsigned)((npairs + x.tl.LP - 1) / x.tl.LP);
}

template <int TN>
static int xnl_t(const Xform &x, double2 *X2, Ctl *ctl) {
  const int lds = 6 * x.tl.LP * x.g.N * (int)sizeof(double2);
  allow_lds(k_xnl<TN>, lds);
  hipLaunchKernelGGL((k_xnl<TN>), dim3(xgrid(x)), dim3(NSL_NT), lds, x.st, X2, x.g, x.hp.P, x.tl.LP, ctl);
  LAUNCH_CHECK();
  return NSL_OK;
}
// *nred: the number of maxima pairs left in pmax for the next y-forward pass to combine (0: already in ctl)
static int xnl(const Xform &x, double2 *X2, double *pmax, Ctl *ctl, int *nred, const Batch &bt = Batch()) {
  *nred = 0;
  if (x.fast) return fast_dispatch_xnl(x, X2, pmax, ctl, bt, nred);
  TRY(need_single(bt));
  return x.tl.TNx == 4096 ? xnl_t<4096>(x, X2, ctl) : xnl_t<8192>(x, X2, ctl);
}

template <int TN>
static int xfwd_t(const Xform &x, const double *phys, double2 *X, int ncomp) {
  const int lds = std::max(ncomp * x.tl.LP * x.g.N, 1) * (int)sizeof(double2);
  allow_lds(k_xfwd<TN>, lds);
  hipLaunchKernelGGL((k_xfwd<TN>), dim3(xgrid(x)), dim3(NSL_NT), lds, x.st, phys, X, x.g, x.hp.P, x.tl.LP, ncomp);
  LAUNCH_CHECK();
  return NSL_OK;
}
static int xfwd(const Xform &x, const double *phys, double2 *X, int ncomp) {
  return x.tl.TNx == 4096 ? xfwd_t<4096>(x, phys, X, ncomp) : xfwd_t<8192>(x, phys, X, ncomp);
}

template <int TN>
static int xinv_t(const Xform &x, const double2 *X, double *phys, int ncomp) {
  const int lds = std::max(ncomp * x.tl.LP * x.g.N, 1) * (int)sizeof(double2);
  allow_lds(k_xinv<TN>, lds);
  hipLaunchKernelGGL((k_xinv<TN>), dim3(xgrid(x)), dim3(NSL_NT), lds, x.st, X, phys, x.g, x.hp.P, x.tl.LP, ncomp);
  LAUNCH_CHECK();
  return NSL_OK;
}
static int xinv(const Xform &x, const double2 *X, double *phys, int ncomp) {
  return x.tl.TNx == 4096 ? xinv_t<4096>(x, X, phys, ncomp) : xinv_t<8192>(x, X, phys, ncomp);
}

static int xform_init(Xform &x, const Geo &g, hipStream_t st) {
  x.g = g;
  x.st = st;
  if (!tiling_for(g.N, x.tl)) return fail(NSL_ERR_ARG, "grid size " + std::to_string(g.N) + " exceeds the LDS tiling of this library (N <= 1024)");
  TRY(plan_create(g.N, x.hp));
  // the size-specialised kernels hard-code the 2/3-truncated geometry; untruncated transforms use the generic ones
  if (g.by != g.N && fast_setup(x)) {
    std::vector<double2> tw(g.N);
    for (int k = 0; k < g.N; k++) {
      const double a = TWO_PI * k / g.N;
      tw[k] = make_double2(cos(a), sin(a));
    }
    HIP_TRY(hipMalloc(&x.twN, g.N * sizeof(double2)));
    HIP_TRY(hipMemcpy(x.twN, tw.data(), g.N * sizeof(double2), hipMemcpyHostToDevice));
    std::vector<float2> tw32(g.N);
    for (int k = 0; k < g.N; k++) tw32[k] = make_float2((float)tw[k].x, (float)tw[k].y);
    HIP_TRY(hipMalloc(&x.twN32, g.N * sizeof(float2)));
    HIP_TRY(hipMemcpy(x.twN32, tw32.data(), g.N * sizeof(float2), hipMemcpyHostToDevice));
  }
  return NSL_OK;
}
static void xform_free(Xform &x) {
  if (x.hp.twd) (void)hipFree(x.hp.twd);
  if (x.twN) (void)hipFree(x.twN);
  if (x.twN32) (void)hipFree(x.twN32);
  x.hp.twd = nullptr;
  x.twN = nullptr;
  x.twN32 = nullptr;
}

#include "nsl_fastdisp.hpp"
#include "nsl_tune.hpp"
#include "nsl_fp32.hpp"

// ------------------------------------------------------------------------------------------------ solver
#define NSL_RECCAP 64

// Device buffers of one solve, carved from one allocation (a "slot"); a pool places its slots SB bytes apart.
struct SlotLayout {
  size_t S, T, ACC, NL, X1, X2, partials, pmax, ctl, recs, bytes;
};

// esz: bytes per stored complex value (16 for FP64 solves, 8 for FP32); an FP32 solve keeps no NL (its update is
// always the fused z-forward, which holds the nonlinear term on chip).
static SlotLayout slot_layout(const Geo &g, int nblk, size_t esz = sizeof(double2), bool with_nl = true) {
  SlotLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t D3 = (size_t)3 * g.D * esz;
  L.S = take(D3); L.T = take(D3); L.ACC = take(D3); L.NL = with_nl ? take(D3) : 0;
  L.X1 = take((size_t)6 * g.N * g.by * g.hp * esz);
  L.X2 = take((size_t)6 * g.N * g.N * g.hp * esz);
  // per-block sums of the spectral sweep (nblk blocks) or of the fused z-forward (at most by * ceil(hx / 4))
  L.partials = take((size_t)std::max(nblk, g.by * ((g.hx + 3) / 4)) * 8 * sizeof(double));
  L.pmax = take((size_t)2 * g.N * g.N * sizeof(double));   // one (max |omega|^2, max |u|^2) per x-pass workgroup
  L.ctl = take(sizeof(Ctl));
  L.recs = take(NSL_RECCAP * sizeof(StepRecDev));
  L.bytes = (off + 4095) & ~(size_t)4095;
  return L;
}

static int nblk_of(const Geo &g) { return (int)((g.D + (long long)NSL_NT * NSL_MPT - 1) / ((long long)NSL_NT * NSL_MPT)); }

struct nsl_solver {
  int N = 0, ic = 0;
  nsl_pool *pool = nullptr;   // a slot view: its pool and slot index
  int slot = -1;
  int dev = 0;                // the GPU the solver's memory and stream live on
  double nu = 0, cfl = 0, dtFixed = 0, icp[4] = {0, 0, 0, 0};
  Xform x;
  hipStream_t st = 0;
  bool view = false;        // a slot of a pool: stream, transform set and slot memory belong to the pool
  Batch bt;                 // {1, 0}, or {slots, SB} for the launch object of a pool
  char *block = nullptr;    // the slot allocation (owned unless view)
  double2 *S = nullptr, *T = nullptr, *ACC = nullptr, *NL = nullptr, *X1 = nullptr, *X2 = nullptr;
  double *partials = nullptr, *pmax = nullptr;
  int nblk = 0;
  Ctl *ctl = nullptr;
  StepRecDev *recs = nullptr;
  StepRecDev *hrecs = nullptr;   // pinned host copy
  Ctl *hctl = nullptr;           // pinned host copy
  int64_t bytes = 0;
  // host mirror of st / last
  nsl_rhsdiag last;
  double t = 0, dt = 0, E0 = 0, Z0 = 0, divMax = 0;
  int64_t step = 0;
  // diagnose buffers (allocated on first use)
  double *phys = nullptr;   // [15][N^3]
  double *dpart = nullptr;  // diagnose partials
  int diag_valid = 0;
  hipEvent_t evfence = nullptr;   // a pool's launch object: orders a group launch after the pool's own stream
  // FP32 solve (nsl_create_fp32): the stepped buffers in complex64 (block); S..X2 above then point into `aux`, an
  // FP64 set that exists only while a host view (initial field, state in or out, diagnostics) runs
  int fp32 = 0;
  float2 *S32 = nullptr, *T32 = nullptr, *ACC32 = nullptr, *X1_32 = nullptr, *X2_32 = nullptr;
  char *aux = nullptr;
  // optional per-kernel timing with HIP events on the solver's stream (nsl_profile)
  int prof = 0;
  std::vector<hipEvent_t> evs;
  std::vector<int> evcls;
  size_t evn = 0;
  double prof_ms[NSL_PROF_CLASSES] = {0};
  int64_t prof_cnt[NSL_PROF_CLASSES] = {0};
};

// Event pair around one launch when profiling is on.
static int prof_begin(nsl_solver *s, int cls) {
  if (!s->prof) return NSL_OK;
  if (s->evn + 2 > s->evs.size()) {
    for (int i = 0; i < 2; i++) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      s->evs.push_back(e);
    }
    s->evcls.resize(s->evs.size() / 2);
  }
  s->evcls[s->evn / 2] = cls;
  HIP_TRY(hipEventRecord(s->evs[s->evn], s->st));
  return NSL_OK;
}
static int prof_end(nsl_solver *s) {
  if (!s->prof) return NSL_OK;
  HIP_TRY(hipEventRecord(s->evs[s->evn + 1], s->st));
  s->evn += 2;
  return NSL_OK;
}
// After a stream synchronisation: fold the recorded pairs into the per-class totals.
static int prof_collect(nsl_solver *s) {
  for (size_t i = 0; i + 1 < s->evn; i += 2) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, s->evs[i], s->evs[i + 1]));
    const int c = s->evcls[i / 2];
    s->prof_ms[c] += ms;
    s->prof_cnt[c]++;
  }
  s->evn = 0;
  return NSL_OK;
}

template <typename T>
static int dalloc(nsl_solver *s, T **p, size_t count) {
  HIP_TRY(hipMalloc((void **)p, count * sizeof(T)));
  s->bytes += (int64_t)(count * sizeof(T));
  return NSL_OK;
}

static void bind_slot(nsl_solver *s, char *base, const SlotLayout &L) {
  s->block = base;
  if (s->fp32) {
    s->S32 = (float2 *)(base + L.S); s->T32 = (float2 *)(base + L.T); s->ACC32 = (float2 *)(base + L.ACC);
    s->X1_32 = (float2 *)(base + L.X1); s->X2_32 = (float2 *)(base + L.X2);
  } else {
    s->S = (double2 *)(base + L.S); s->T = (double2 *)(base + L.T); s->ACC = (double2 *)(base + L.ACC);
    s->NL = (double2 *)(base + L.NL); s->X1 = (double2 *)(base + L.X1); s->X2 = (double2 *)(base + L.X2);
  }
  s->partials = (double *)(base + L.partials); s->pmax = (double *)(base + L.pmax);
  s->ctl = (Ctl *)(base + L.ctl); s->recs = (StepRecDev *)(base + L.recs);
}


static int prof_collect(nsl_solver *s);
static int pull_ctl(nsl_solver *s) {
  HIP_TRY(hipMemcpyAsync(s->hctl, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return prof_collect(s);
}
static int push_ctl(nsl_solver *s) {
  HIP_TRY(hipMemcpyAsync(s->ctl, s->hctl, sizeof(Ctl), hipMemcpyHostToDevice, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}
static void diag_from(const double *d, nsl_rhsdiag &o) {
  o.E = d[0]; o.Z = d[1]; o.P = d[2]; o.eps = d[3]; o.Tnl = d[4]; o.Pspec = d[5]; o.omMax = d[6]; o.uMax = d[7];
}

// Rows of partial sums a fused update leaves for the next launch to book (Ctl::pend), or 0 when the solver's rhs
// books them at once (the generic and two-kernel paths).
static int pending_rows(const nsl_solver *s) { return s->fp32 || fast_has_zu(s->x) ? fast_zu_blocks(s->x.g) : 0; }
// The booking of the rhs before this one, when its fused update left it pending: done by the z-inverse of this rhs.
static Fin pending_fin(const nsl_solver *s, int prev_fmode) {
  Fin f;
  f.partials = s->partials; f.nfin = pending_rows(s); f.fmode = prev_fmode;
  return f;
}
// Books whatever a fused update left pending (mode 4: the rhs of the current state; with rec, the end of a step):
// after the last step enqueued, and after a rhs outside a step. Nothing pending, nothing done.
static int flush_pending(nsl_solver *s, StepRecDev *rec) {
  const int nfin = pending_rows(s);
  if (nfin == 0) return NSL_OK;
  hipLaunchKernelGGL(k_finalize, dim3(1, 1, s->bt.nz), dim3(NSL_NT), 0, s->st, s->partials, nfin, 4, rec, s->ctl,
                     s->bt.sb, 1);
  LAUNCH_CHECK();
  return NSL_OK;
}

// One rhs of an FP32 solve's S32 (onS) or T32, with the fused stage-`stage` update (its sums booked later, as in
// rhs_enqueue); fin: the previous rhs's booking.
static int rhs_enqueue32(nsl_solver *s, bool onS, int stage, const Fin &fin) {
  const Xform &x = s->x;
  const float2 *U = onS ? s->S32 : s->T32;
  TRY(prof_begin(s, NSL_PROF_ZINV)); TRY(zinv32(x, U, s->X1_32, s->ctl, fin)); TRY(prof_end(s));
  TRY(prof_begin(s, NSL_PROF_YINV)); TRY(yinv32(x, s->X1_32, s->X2_32, s->ctl)); TRY(prof_end(s));
  int nred = 0;
  TRY(prof_begin(s, NSL_PROF_XNL)); TRY(xnl32(x, s->X2_32, s->pmax, s->ctl, &nred)); TRY(prof_end(s));
  TRY(prof_begin(s, NSL_PROF_YFWD)); TRY(yfwd32(x, s->X2_32, s->X1_32, s->ctl, s->pmax, nred)); TRY(prof_end(s));
  int nsum = 0;
  TRY(prof_begin(s, stage == 1 ? NSL_PROF_UPDATE1 : (stage == 4 ? NSL_PROF_UPDATE4 : NSL_PROF_UPDATE23)));
  TRY(zu32(x, stage, s->X1_32, U, s->S32, s->ACC32, s->T32, s->partials, s->ctl, &nsum));
  TRY(prof_end(s));
  return NSL_OK;
}

// One rhs of S (onS) or T with the stage-`stage` update; fmode selects where the sums are booked (diag[fmode]).
// The fused update leaves its sums pending (Ctl::pend) and they are booked by the next launch that follows: the next
// rhs's z-inverse (fin, one extra workgroup), the next step's k_begin_step or flush_pending, instead of a
// single-workgroup kernel after every rhs. The generic and two-kernel paths book them at once (k_finalize).
static int rhs_enqueue(nsl_solver *s, bool onS, int stage, int fmode, StepRecDev *rec, const Fin &fin) {
  if (s->fp32) return rhs_enqueue32(s, onS, stage, fin);
  const Xform &x = s->x;
  const Batch &bt = s->bt;
  double2 *U = onS ? s->S : s->T;
  TRY(prof_begin(s, NSL_PROF_ZINV)); TRY(zinv(x, U, s->X1, 1, 6, 0, s->ctl, bt, fin)); TRY(prof_end(s));
  if (fast_has_plane(x)) {   // y-inverse, x pass and y-forward in one pass (timed as the x pass)
    TRY(prof_begin(s, NSL_PROF_XNL)); TRY(fast_dispatch_plane(x, s->X1, s->ctl, bt)); TRY(prof_end(s));
  } else {
    TRY(prof_begin(s, NSL_PROF_YINV)); TRY(yinv(x, s->X1, s->X2, 6, s->ctl, bt)); TRY(prof_end(s));
    int nred = 0;
    TRY(prof_begin(s, NSL_PROF_XNL)); TRY(xnl(x, s->X2, s->pmax, s->ctl, &nred, bt)); TRY(prof_end(s));
    TRY(prof_begin(s, NSL_PROF_YFWD)); TRY(yfwd(x, s->X2, s->X1, 3, s->ctl, bt, s->pmax, nred)); TRY(prof_end(s));
  }
  if (fast_has_zu(x)) {
    // z-forward and update in one pass: the nonlinear term stays on chip (timed as the update of this stage)
    TRY(prof_begin(s, stage == 1 ? NSL_PROF_UPDATE1 : (stage == 4 ? NSL_PROF_UPDATE4 : NSL_PROF_UPDATE23)));
    TRY(fast_dispatch_zu(x, stage, s->X1, U, s->S, s->ACC, s->T, s->partials, s->ctl, bt));
    TRY(prof_end(s));
    return NSL_OK;
  }
  TRY(prof_begin(s, NSL_PROF_ZFWD)); TRY(zfwd(x, s->X1, s->NL, 3, s->ctl, bt)); TRY(prof_end(s));
  dim3 g((unsigned)s->nblk, 1, bt.nz), b(NSL_NT);
  TRY(prof_begin(s, stage == 1 ? NSL_PROF_UPDATE1 : (stage == 4 ? NSL_PROF_UPDATE4 : NSL_PROF_UPDATE23)));
  switch (stage) {
    case 1: hipLaunchKernelGGL(k_update<1>, g, b, 0, s->st, s->NL, U, s->S, s->ACC, s->T, x.g, s->partials, s->ctl, bt.sb); break;
    case 2: hipLaunchKernelGGL(k_update<2>, g, b, 0, s->st, s->NL, U, s->S, s->ACC, s->T, x.g, s->partials, s->ctl, bt.sb); break;
    case 3: hipLaunchKernelGGL(k_update<3>, g, b, 0, s->st, s->NL, U, s->S, s->ACC, s->T, x.g, s->partials, s->ctl, bt.sb); break;
    default: hipLaunchKernelGGL(k_update<4>, g, b, 0, s->st, s->NL, U, s->S, s->ACC, s->T, x.g, s->partials, s->ctl, bt.sb); break;
  }
  LAUNCH_CHECK();
  TRY(prof_end(s));
  TRY(prof_begin(s, NSL_PROF_FINALIZE));
  hipLaunchKernelGGL(k_finalize, dim3(1, 1, bt.nz), dim3(NSL_NT), 0, s->st, s->partials, s->nblk, fmode, rec, s->ctl, bt.sb, 0);
  LAUNCH_CHECK();
  TRY(prof_end(s));
  return NSL_OK;
}

// One RK4 step (nslab.js:312-333). Stage 1's rhs is the previous step's last rhs (identical input), already
// booked in diag[4] with ACC and T prepared; the step ends with the rhs of the new state. k_begin_step first books
// the previous step's end if it is pending (budgets, t, step, one record).
static int step_enqueue(nsl_solver *s) {
  hipLaunchKernelGGL(k_begin_step, dim3(1, 1, s->bt.nz), dim3(NSL_NT), 0, s->st, s->ctl, (const double *)s->partials,
                     pending_rows(s), s->recs, s->bt.sb);
  LAUNCH_CHECK();
  TRY(rhs_enqueue(s, false, 2, 1, nullptr, Fin()));
  TRY(rhs_enqueue(s, false, 3, 2, nullptr, pending_fin(s, 1)));
  TRY(rhs_enqueue(s, false, 4, 3, nullptr, pending_fin(s, 2)));
  TRY(rhs_enqueue(s, true, 1, 4, s->recs, pending_fin(s, 3)));
  return NSL_OK;
}

// rhs of the current state outside a step (finishIC, set_state, diagnose): refreshes last, ACC and T.
static int rhs_current(nsl_solver *s) {
  hipLaunchKernelGGL(k_set_active, dim3(1), dim3(1), 0, s->st, s->ctl, 1, (size_t)0);
  LAUNCH_CHECK();
  TRY(rhs_enqueue(s, true, 1, 4, nullptr, Fin()));
  TRY(flush_pending(s, nullptr));
  TRY(pull_ctl(s));
  diag_from(s->hctl->diag[4], s->last);
  return NSL_OK;
}

static unsigned sweep_grid(long long n) {
  long long b = (n + NSL_NT - 1) / NSL_NT;
  return (unsigned)std::min<long long>(std::max<long long>(b, 1), 65536);
}

// Forward transform of the 3 physical fields in X2 (as [3][N^3] reals) into compact spectral OUT.
static int forward3_from_phys(nsl_solver *s, double2 *OUT) {
  const Xform &x = s->x;
  double *phys = (double *)s->X2;
  TRY(xfwd(x, phys, s->X1, 3));
  TRY(yfwd(x, s->X1, s->X2, 3, nullptr));
  TRY(zfwd(x, s->X2, OUT, 3, nullptr));
  return NSL_OK;
}

static int norms_dev(nsl_solver *s, const double2 *U, double *out3_dev) {
  hipLaunchKernelGGL(k_norms, dim3((unsigned)s->nblk), dim3(NSL_NT), 0, s->st, U, s->x.g, s->partials);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_norms_finalize, dim3(1), dim3(NSL_NT), 0, s->st, s->partials, s->nblk, out3_dev);
  LAUNCH_CHECK();
  return NSL_OK;
}

// ------------------------------------------------------------------------------------------------ FP32 solves
// Widen the complex64 state to its FP64 view, or round the FP64 view into the complex64 state (n complex values).
__global__ __launch_bounds__(NSL_NT) void k_widen(const float2 *__restrict__ a, double2 *__restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * NSL_NT + threadIdx.x; i < n; i += (long long)gridDim.x * NSL_NT)
    b[i] = make_double2((double)a[i].x, (double)a[i].y);
}
__global__ __launch_bounds__(NSL_NT) void k_narrow(const double2 *__restrict__ a, float2 *__restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * NSL_NT + threadIdx.x; i < n; i += (long long)gridDim.x * NSL_NT)
    b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// The FP64 buffers of an FP32 solve's host views, allocated for the view and freed after it: S (the state,
// widened when `widen`), T and NL (scratch of the initial fields), X1 and X2 (transform scratch of the initial
// fields and the diagnostics). A no-op for FP64 solves, whose S..X2 are their own.
static int aux_acquire(nsl_solver *s, bool widen) {
  if (!s->fp32) return NSL_OK;
  const Geo &g = s->x.g;
  if (!s->aux) {
    size_t off = 0;
    auto take = [&](size_t b) {
      const size_t o = off;
      off += (b + 255) & ~(size_t)255;
      return o;
    };
    const size_t D3 = (size_t)3 * g.D * sizeof(double2);
    const size_t oS = take(D3), oT = take(D3), oNL = take(D3);
    const size_t oX1 = take((size_t)6 * g.N * g.by * g.hp * sizeof(double2));
    const size_t oX2 = take((size_t)6 * g.N * g.N * g.hp * sizeof(double2));
    char *p = nullptr;
    const hipError_t e = hipMalloc((void **)&p, off);
    if (e != hipSuccess)
      return fail(NSL_ERR_HIP, "FP64 view of an FP32 solve: hipMalloc of " + std::to_string(off) + " bytes: " + hipGetErrorString(e));
    s->aux = p;
    s->S = (double2 *)(p + oS); s->T = (double2 *)(p + oT); s->NL = (double2 *)(p + oNL);
    s->X1 = (double2 *)(p + oX1); s->X2 = (double2 *)(p + oX2);
    HIP_TRY(hipMemsetAsync(p, 0, off, s->st));   // padding slots of S read as zeros
  }
  if (widen) {
    hipLaunchKernelGGL(k_widen, dim3(sweep_grid(3 * g.D)), dim3(NSL_NT), 0, s->st, (const float2 *)s->S32, s->S, 3 * g.D);
    LAUNCH_CHECK();
  }
  return NSL_OK;
}
// Round the FP64 view of the state into the FP32 state.
static int aux_narrow(nsl_solver *s) {
  const Geo &g = s->x.g;
  hipLaunchKernelGGL(k_narrow, dim3(sweep_grid(3 * g.D)), dim3(NSL_NT), 0, s->st, (const double2 *)s->S, s->S32, 3 * g.D);
  LAUNCH_CHECK();
  return NSL_OK;
}
static void aux_release(nsl_solver *s) {
  if (!s->fp32 || !s->aux) return;
  (void)hipStreamSynchronize(s->st);
  (void)hipFree(s->aux);
  s->aux = nullptr;
  s->S = s->T = s->NL = s->X1 = s->X2 = nullptr;
}
// The FP64 view for the duration of one host call.
struct AuxView {
  nsl_solver *s;
  int rc;
  AuxView(nsl_solver *s_, bool widen) : s(s_), rc(aux_acquire(s_, widen)) {}
  ~AuxView() { aux_release(s); }
};

static inline int full_of_h(int c, const Geo &g) { return c < g.hy ? c : c + g.N - g.by; }
static inline int compact_of_h(int f, const Geo &g) {
  return f < g.hy ? f : (f >= g.N - g.by + g.hy ? f - (g.N - g.by) : -1);
}
static inline double kw_h(int f, int N) { return (double)(f <= N / 2 ? f : f - N); }

// Random solenoidal field (nslab.js:273-286): the draws of every mode on the device (k_ic_random, each mode from its
// own position in the reference's draw sequence), then projected and scaled to E0.
static int ic_random(nsl_solver *s, double2 *dst, double seed, double k0, double E0) {
  const Geo &g = s->x.g;
  HIP_TRY(hipMemsetAsync(dst, 0, (size_t)3 * g.D * sizeof(double2), s->st));
  const long long nm = (long long)g.by * g.by * g.hx;
  hipLaunchKernelGGL(k_ic_random, dim3((unsigned)((nm + NSL_NT - 1) / NSL_NT)), dim3(NSL_NT), 0, s->st, dst, g,
                     (unsigned)(int32_t)seed, k0);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_project, dim3(sweep_grid(g.D)), dim3(NSL_NT), 0, s->st, dst, g);
  LAUNCH_CHECK();
  double *nrm = (double *)((char *)s->ctl + offsetof(Ctl, nrm));   // device scratch for the three norms
  TRY(norms_dev(s, dst, nrm));
  hipLaunchKernelGGL(k_scale_energy, dim3(sweep_grid(3 * g.D)), dim3(NSL_NT), 0, s->st, dst, 3 * g.D, (const double *)nrm, E0);
  LAUNCH_CHECK();
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}

static int set_ic(nsl_solver *s, int ic, const double *icp) {
  if (ic < 0 || ic > NSL_IC_TUBESZ) return fail(NSL_ERR_ARG, "unknown initial condition " + std::to_string(ic));
  s->ic = ic;
  for (int q = 0; q < 4; q++) s->icp[q] = icp ? icp[q] : 0;
  const Geo &g = s->x.g;
  const int N = g.N;
  AuxView av(s, false);   // FP32 solves: the field is built in FP64, then rounded into the FP32 state
  TRY(av.rc);
  if (ic == NSL_IC_RANDOM) {
    TRY(ic_random(s, s->S, s->icp[0], s->icp[1], s->icp[2]));
  } else {
    double *phys = (double *)s->X2;
    hipLaunchKernelGGL(k_ic_phys, dim3(sweep_grid((long long)N * N * N)), dim3(NSL_NT), 0, s->st, phys, N, ic == NSL_IC_TUBES ? 3 : ic == NSL_IC_TUBESZ ? 4 : ic,
                       s->icp[0], s->icp[1], s->icp[2], s->icp[3]);
    LAUNCH_CHECK();
    if (ic == NSL_IC_TUBES || ic == NSL_IC_TUBESZ) {
      TRY(forward3_from_phys(s, s->NL));
      hipLaunchKernelGGL(k_tubes_finish, dim3(sweep_grid(g.D)), dim3(NSL_NT), 0, s->st, s->NL, s->S, g);
      LAUNCH_CHECK();
      if (ic == NSL_IC_TUBESZ && s->icp[3] > 0) {
        // seeded random solenoidal perturbation (k^4 exp(-2 (k/4)^2), seed 1), rms velocity icp[3] per component:
        // energy 3/2 icp[3]^2
        TRY(ic_random(s, s->T, 1.0, 4.0, 1.5 * s->icp[3] * s->icp[3]));
        hipLaunchKernelGGL(k_axpy, dim3(sweep_grid(3 * g.D)), dim3(NSL_NT), 0, s->st, s->S, (const double2 *)s->T, 3 * g.D);
        LAUNCH_CHECK();
      }
    } else {
      TRY(forward3_from_phys(s, s->S));
      hipLaunchKernelGGL(k_project, dim3(sweep_grid(g.D)), dim3(NSL_NT), 0, s->st, s->S, g);
      LAUNCH_CHECK();
    }
  }
  if (s->fp32) {
    TRY(aux_narrow(s));
    aux_release(s);
  }
  // finishIC (nslab.js:293-299): counters reset, one rhs, E0/Z0 from it
  s->t = 0; s->step = 0; s->dt = 0; s->divMax = 0; s->diag_valid = 0;
  TRY(pull_ctl(s));
  Ctl *c = s->hctl;
  c->t = 0; c->dt = 0; c->step = 0; c->om2 = 0; c->u2 = 0; c->nDone = 0; c->nTarget = 0; c->nChunk = 0;
  c->E0ref = 1.0;
  TRY(push_ctl(s));
  TRY(rhs_current(s));
  s->E0 = s->last.E; s->Z0 = s->last.Z;
  c->E0ref = s->E0;
  TRY(push_ctl(s));
  return NSL_OK;
}

static void solver_free(nsl_solver *s) {
  if (!s) return;
  if (s->phys) (void)hipFree(s->phys);
  if (s->dpart) (void)hipFree(s->dpart);
  if (s->hrecs) (void)hipHostFree(s->hrecs);
  if (s->hctl) (void)hipHostFree(s->hctl);
  if (s->aux) (void)hipFree(s->aux);
  for (auto e : s->evs) (void)hipEventDestroy(e);
  if (s->evfence) (void)hipEventDestroy(s->evfence);
  if (!s->view) {
    if (s->block) (void)hipFree(s->block);
    xform_free(s->x);
    if (s->st) (void)hipStreamDestroy(s->st);
  }
  delete s;
}

// Host-side state of a solve whose device buffers are bound: pinned mirrors and the initial control block.
static int solver_host_init(nsl_solver *s) {
  if (hipHostMalloc((void **)&s->hrecs, NSL_RECCAP * sizeof(StepRecDev)) != hipSuccess ||
      hipHostMalloc((void **)&s->hctl, sizeof(Ctl)) != hipSuccess)
    return fail(NSL_ERR_HIP, "hipHostMalloc failed");
  memset(s->hctl, 0, sizeof(Ctl));
  Ctl *c = s->hctl;
  c->N = s->N; c->kc = s->N / 3; c->nu = s->nu; c->cfl = s->cfl; c->dtFixed = s->dtFixed; c->E0ref = 1.0;
  return push_ctl(s);
}

// fp32: complex64 storage of every stepped buffer (nslab_gpu.py:56-65), for the sizes with an FP32 path.
static int create_impl(const nsl_opts *o, nsl_solver **out, bool fp32 = false) {
  if (!o || !out) return fail(NSL_ERR_ARG, "null argument");
  *out = nullptr;
  const int N = o->N;
  std::vector<int> f;
  if (factorize_host(N, f) < 0) return fail(NSL_ERR_ARG, bad_n_msg(N));
  if (N % 2) return fail(NSL_ERR_ARG, "grid size must be even for the real-to-complex transforms (got " + std::to_string(N) + ")");
  if (o->ic < 0 || o->ic > NSL_IC_TUBESZ) return fail(NSL_ERR_ARG, "unknown initial condition " + std::to_string(o->ic));
  if (fp32 && !fp32_has(N))
    return fail(NSL_ERR_ARG, "FP32 storage is compiled for N = 32, 48, 64, 96, 128, 192, 256, 384, 512, 1024 (got " + std::to_string(N) + ")");
  nsl_solver *s = new nsl_solver();
  s->N = N; s->nu = o->nu; s->cfl = o->cfl; s->dtFixed = o->dt; s->fp32 = fp32 ? 1 : 0;
  if (hipGetDevice(&s->dev) != hipSuccess) { delete s; return fail(NSL_ERR_HIP, "hipGetDevice failed"); }
  hipError_t e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) { delete s; return fail(NSL_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
  int r = xform_init(s->x, geo_truncated(N), s->st);
  if (r != NSL_OK) { solver_free(s); return r; }
  s->nblk = nblk_of(s->x.g);
  if (fp32 && !s->x.fast) { solver_free(s); return fail(NSL_ERR_ARG, "FP32 storage needs the size-specialised kernels"); }
  const SlotLayout L = fp32 ? slot_layout(s->x.g, s->nblk, sizeof(float2), false) : slot_layout(s->x.g, s->nblk);
  char *base = nullptr;
  if ((r = dalloc(s, &base, L.bytes)) != NSL_OK) { solver_free(s); return r; }
  bind_slot(s, base, L);
  // the padding slots of every row stay zero for the life of the solver. On the solver's stream: a non-blocking
  // stream does not wait for work on the null stream, so a plain hipMemset could still be clearing the block
  // (control block included) while the set-up kernels run.
  if (hipMemsetAsync(base, 0, L.bytes, s->st) != hipSuccess) { solver_free(s); return fail(NSL_ERR_HIP, "hipMemset failed"); }
  if ((r = solver_host_init(s)) != NSL_OK) { solver_free(s); return r; }
  if ((r = set_ic(s, o->ic, o->icp)) != NSL_OK) { solver_free(s); return r; }
  *out = s;
  return NSL_OK;
}

static int run_impl(nsl_solver *s, int64_t n, int has_tend, double tend, nsl_steprec *recs, int64_t *n_done) {
  if (n_done) *n_done = 0;
  if (n <= 0) return NSL_OK;
  TRY(pull_ctl(s));
  Ctl *c = s->hctl;
  c->nDone = 0; c->nTarget = n; c->hasTEnd = has_tend ? 1 : 0; c->tEnd = tend; c->nChunk = 0;
  c->t = s->t; c->step = s->step; c->E0ref = s->E0;
  TRY(push_ctl(s));
  int64_t enq = 0, done = 0;
  while (enq < n) {
    const int chunk = (int)std::min<int64_t>(NSL_RECCAP, n - enq);
    hipLaunchKernelGGL(k_chunk_begin, dim3(1), dim3(1), 0, s->st, s->ctl, (size_t)0);
    LAUNCH_CHECK();
    for (int i = 0; i < chunk; i++) TRY(step_enqueue(s));
    TRY(flush_pending(s, s->recs));   // the last step's end
    HIP_TRY(hipMemcpyAsync(s->hrecs, s->recs, chunk * sizeof(StepRecDev), hipMemcpyDeviceToHost, s->st));
    TRY(pull_ctl(s));
    const int got = (int)c->nChunk;
    for (int i = 0; i < got; i++) {
      const StepRecDev &R = s->hrecs[i];
      if (recs) {
        nsl_steprec &o = recs[done + i];
        o.t = R.t; o.dt = R.dt; o.step = R.step; o.ebal = R.ebal; o.zbal = R.zbal;
        diag_from(R.last, o.last);
        diag_from(R.d0, o.d0);
      }
    }
    done += got;
    enq += chunk;
    if (got < chunk) break;
  }
  s->t = c->t; s->step = c->step; if (done > 0) s->dt = c->dt;
  diag_from(c->diag[4], s->last);
  if (done > 0) s->diag_valid = 0;
  if (n_done) *n_done = done;
  return NSL_OK;
}

// ------------------------------------------------------------------------------------------------ pools
// A pool holds `slots` solves of one N in one allocation, slot i at base + i*SB. One launch of every pass covers
// all slots (gridDim.z = slots); each slot's control block decides whether its workgroups work or return, so
// solves join and leave between chunks without changing the launches, and a slot's arithmetic (launch shape per
// slot, reduction order) is the same as a single solve's whatever the other slots hold. One RK4 step of the pool
// is captured once as a hipGraph and replayed.
struct PoolStage {
  Ctl ctl;
  StepRecDev recs[NSL_RECCAP];
};

// Copies each slot's control block and step records into a contiguous staging array (one copy to the host).
__global__ void k_pool_gather(const char *__restrict__ base, size_t sb, size_t offCtl, size_t offRecs, int nrec,
                              PoolStage *__restrict__ out) {
  const char *slot = base + (size_t)blockIdx.x * sb;
  const double *c = (const double *)(slot + offCtl);
  double *oc = (double *)&out[blockIdx.x].ctl;
  for (int i = threadIdx.x; i < (int)(sizeof(Ctl) / sizeof(double)); i += blockDim.x) oc[i] = c[i];
  const double *r = (const double *)(slot + offRecs);
  double *orr = (double *)out[blockIdx.x].recs;
  const int nr = nrec * (int)(sizeof(StepRecDev) / sizeof(double));
  for (int i = threadIdx.x; i < nr; i += blockDim.x) orr[i] = r[i];
}

struct nsl_pool {
  int N = 0, C = 0, nblk = 0;
  int dev = 0;
  Xform x;
  hipStream_t st = 0;
  SlotLayout L;
  char *base = nullptr;
  nsl_solver *lead = nullptr;            // the launch object: slot 0's buffers, Batch {C, SB}
  std::vector<nsl_solver *> slots;       // one view per slot
  std::vector<int> used;
  // per-slot generation, bumped by admit / set_target / release; launch_gen is its value at the last launch, so
  // collect can tell slots changed while a chunk was in flight (their staged control block predates the change)
  std::vector<uint64_t> gen, launch_gen;
  PoolStage *dstage = nullptr, *hstage = nullptr;
  int lastChunk = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graphed = 0;                       // 1: steps replay the captured graph; 0: eager launches
  int64_t bytes = 0;
};

static int64_t slot_bytes_of(int N) {
  const Geo g = geo_truncated(N);
  return (int64_t)slot_layout(g, nblk_of(g)).bytes;
}

static void pool_free(nsl_pool *p) {
  if (!p) return;
  if (p->st) (void)hipStreamSynchronize(p->st);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (nsl_solver *v : p->slots) solver_free(v);
  solver_free(p->lead);
  if (p->dstage) (void)hipFree(p->dstage);
  if (p->hstage) (void)hipHostFree(p->hstage);
  if (p->base) (void)hipFree(p->base);
  xform_free(p->x);
  if (p->st) (void)hipStreamDestroy(p->st);
  delete p;
}

static nsl_solver *pool_view(nsl_pool *p, char *base, Batch bt, int slot) {
  nsl_solver *v = new nsl_solver();
  v->view = true;
  v->pool = p;
  v->slot = slot;
  v->dev = p->dev;
  v->N = p->N;
  v->x = p->x;
  v->st = p->st;
  v->bt = bt;
  v->nblk = p->nblk;
  bind_slot(v, base, p->L);
  return v;
}

static int pool_create_impl(int N, int C, nsl_pool **out) {
  if (!out) return fail(NSL_ERR_ARG, "null argument");
  *out = nullptr;
  std::vector<int> f;
  if (factorize_host(N, f) < 0) return fail(NSL_ERR_ARG, bad_n_msg(N));
  if (C < 1 || C > 65535) return fail(NSL_ERR_ARG, "slots must be in 1..65535");
  nsl_pool *p = new nsl_pool();
  p->N = N; p->C = C;
  if (hipGetDevice(&p->dev) != hipSuccess) { delete p; return fail(NSL_ERR_HIP, "hipGetDevice failed"); }
  hipError_t e = hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking);
  if (e != hipSuccess) { delete p; return fail(NSL_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
  int r = xform_init(p->x, geo_truncated(N), p->st);
  if (r == NSL_OK && !p->x.fast) r = fail(NSL_ERR_ARG, "pools need a grid size with compiled kernels (16..1024, 2^a or 2^a*3)");
  if (r != NSL_OK) { pool_free(p); return r; }
  p->nblk = nblk_of(p->x.g);
  p->L = slot_layout(p->x.g, p->nblk);
  const size_t total = (size_t)C * p->L.bytes;
  if (hipMalloc((void **)&p->base, total) != hipSuccess) {
    (void)hipGetLastError();
    pool_free(p);
    return fail(NSL_ERR_HIP, "hipMalloc of " + std::to_string(total) + " bytes for " + std::to_string(C) + " slots failed");
  }
  p->bytes = (int64_t)total;
  // zero: padding stays zero for the pool's life, and every control block starts idle (nTarget = 0)
  if (hipMemsetAsync(p->base, 0, total, p->st) != hipSuccess) { pool_free(p); return fail(NSL_ERR_HIP, "hipMemset failed"); }
  if (hipMalloc((void **)&p->dstage, (size_t)C * sizeof(PoolStage)) != hipSuccess ||
      hipHostMalloc((void **)&p->hstage, (size_t)C * sizeof(PoolStage)) != hipSuccess) {
    pool_free(p);
    return fail(NSL_ERR_HIP, "staging allocation failed");
  }
  p->lead = pool_view(p, p->base, Batch{(unsigned)C, p->L.bytes}, -1);
  if (hipEventCreateWithFlags(&p->lead->evfence, hipEventDisableTiming) != hipSuccess) {
    pool_free(p);
    return fail(NSL_ERR_HIP, "hipEventCreate failed");
  }
  for (int i = 0; i < C; i++) p->slots.push_back(pool_view(p, p->base + (size_t)i * p->L.bytes, Batch(), i));
  p->used.assign(C, 0);
  p->gen.assign(C, 0);
  p->launch_gen.assign(C, 0);
  // one idle step (every slot inactive: each kernel returns at once) sets the launch attributes, then the
  // step is captured
  if ((r = step_enqueue(p->lead)) != NSL_OK) { pool_free(p); return r; }
  HIP_TRY(hipStreamSynchronize(p->st));
  if (hipStreamBeginCapture(p->st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    const int rc = step_enqueue(p->lead);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(p->st, &g);
    if (rc == NSL_OK && ec == hipSuccess && g != nullptr &&
        hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0) == hipSuccess) {
      p->graph = g;
      p->graphed = 1;
    } else if (g != nullptr) {
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipGetLastError();
  *out = p;
  return NSL_OK;
}

// A slot that took part in the chunk still in flight (launched, not yet collected): its control block and state are
// the chunk's until pool_collect books them, so nothing may re-target, step or overwrite it before then. A slot
// admitted or re-targeted after the launch was not part of it (its generation moved on).
static int slot_in_flight(const nsl_solver *s) {
  if (!s->view || !s->pool || s->slot < 0) return NSL_OK;
  const nsl_pool *p = s->pool;
  if (p->lastChunk > 0 && p->gen[s->slot] == p->launch_gen[s->slot])
    return fail(NSL_ERR_STATE, "slot " + std::to_string(s->slot) + " is in the pool's chunk in flight: collect the pool first");
  return NSL_OK;
}

static int pool_slot_check(nsl_pool *p, int slot, bool must_be_used) {
  if (!p) return fail(NSL_ERR_ARG, "null pool");
  if (slot < 0 || slot >= p->C) return fail(NSL_ERR_ARG, "slot out of range");
  if (must_be_used && !p->used[slot]) return fail(NSL_ERR_STATE, "slot is free");
  return NSL_OK;
}

static int pool_admit_impl(nsl_pool *p, const nsl_opts *o, int *slot) {
  if (!p || !o || !slot) return fail(NSL_ERR_ARG, "null argument");
  if (o->N != p->N) return fail(NSL_ERR_ARG, "grid size " + std::to_string(o->N) + " does not match the pool's " + std::to_string(p->N));
  if (o->ic < 0 || o->ic > NSL_IC_TUBESZ) return fail(NSL_ERR_ARG, "unknown initial condition " + std::to_string(o->ic));
  int i = 0;
  while (i < p->C && p->used[i]) i++;
  if (i == p->C) return fail(NSL_ERR_STATE, "pool is full");
  nsl_solver *v = p->slots[i];
  v->nu = o->nu; v->cfl = o->cfl; v->dtFixed = o->dt;
  v->t = 0; v->step = 0; v->dt = 0; v->divMax = 0; v->diag_valid = 0;
  if (!v->hctl) {
    TRY(solver_host_init(v));
  } else {
    memset(v->hctl, 0, sizeof(Ctl));
    Ctl *c = v->hctl;
    c->N = v->N; c->kc = v->N / 3; c->nu = v->nu; c->cfl = v->cfl; c->dtFixed = v->dtFixed; c->E0ref = 1.0;
    TRY(push_ctl(v));
  }
  TRY(set_ic(v, o->ic, o->icp));
  p->used[i] = 1;
  p->gen[i]++;
  *slot = i;
  return NSL_OK;
}

static int pool_release_impl(nsl_pool *p, int slot) {
  TRY(pool_slot_check(p, slot, true));
  nsl_solver *v = p->slots[slot];
  TRY(pull_ctl(v));
  v->hctl->nTarget = v->hctl->nDone;   // idle from the next step on
  TRY(push_ctl(v));
  if (v->phys) { (void)hipFree(v->phys); v->phys = nullptr; }
  if (v->dpart) { (void)hipFree(v->dpart); v->dpart = nullptr; }
  p->used[slot] = 0;
  p->gen[slot]++;
  return NSL_OK;
}

// The next chunk(s) of this slot run up to n steps (and stop at tEnd when has_tend), as run() would.
static int pool_set_target_impl(nsl_pool *p, int slot, int64_t n, int has_tend, double tend) {
  TRY(pool_slot_check(p, slot, true));
  nsl_solver *v = p->slots[slot];
  TRY(slot_in_flight(v));
  TRY(pull_ctl(v));
  Ctl *c = v->hctl;
  c->nDone = 0; c->nTarget = n > 0 ? n : 0; c->hasTEnd = has_tend ? 1 : 0; c->tEnd = tend; c->nChunk = 0;
  c->t = v->t; c->step = v->step; c->E0ref = v->E0;
  p->gen[slot]++;
  return push_ctl(v);
}

// Enqueues up to `steps` (<= NSL_RECCAP) steps of every slot on the pool's stream; returns without waiting.
static int pool_launch_impl(nsl_pool *p, int steps) {
  if (!p) return fail(NSL_ERR_ARG, "null pool");
  if (steps < 1 || steps > NSL_RECCAP) return fail(NSL_ERR_ARG, "steps per launch must be in 1..64");
  hipLaunchKernelGGL(k_chunk_begin, dim3(1, 1, p->C), dim3(1), 0, p->st, p->lead->ctl, p->L.bytes);
  LAUNCH_CHECK();
  for (int i = 0; i < steps; i++) {
    if (p->graphed) HIP_TRY(hipGraphLaunch(p->gexec, p->st));
    else TRY(step_enqueue(p->lead));
  }
  TRY(flush_pending(p->lead, p->lead->recs));   // every slot's last step's end
  hipLaunchKernelGGL(k_pool_gather, dim3(p->C), dim3(256), 0, p->st, (const char *)p->base, p->L.bytes, p->L.ctl,
                     p->L.recs, steps, p->dstage);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(p->hstage, p->dstage, (size_t)p->C * sizeof(PoolStage), hipMemcpyDeviceToHost, p->st));
  p->lastChunk = steps;
  p->launch_gen = p->gen;
  return NSL_OK;
}

// Waits for the launched chunk and books it: per slot, the steps done (n_done[slot]) and their records
// (recs[slot * recs_per_slot + i], if recs); *n_running counts slots with steps left.
static int pool_collect_impl(nsl_pool *p, nsl_steprec *recs, int recs_per_slot, int64_t *n_done, int *n_running) {
  if (!p) return fail(NSL_ERR_ARG, "null pool");
  HIP_TRY(hipStreamSynchronize(p->st));
  int running = 0;
  for (int i = 0; i < p->C; i++) {
    if (n_done) n_done[i] = 0;
    if (!p->used[i]) continue;
    nsl_solver *v = p->slots[i];
    if (p->gen[i] != p->launch_gen[i]) {   // admitted or re-targeted after the launch: nothing of it was staged
      const Ctl &c = *v->hctl;
      if (c.nDone < c.nTarget && !(c.hasTEnd && c.t >= c.tEnd - 1e-12)) running++;
      continue;
    }
    const PoolStage &P = p->hstage[i];
    memcpy(v->hctl, &P.ctl, sizeof(Ctl));
    const Ctl &c = P.ctl;
    const int got = p->lastChunk > 0 ? (int)c.nChunk : 0;
    for (int k = 0; k < got; k++) {
      if (recs && k < recs_per_slot) {
        const StepRecDev &R = P.recs[k];
        nsl_steprec &o = recs[(size_t)i * recs_per_slot + k];
        o.t = R.t; o.dt = R.dt; o.step = R.step; o.ebal = R.ebal; o.zbal = R.zbal;
        diag_from(R.last, o.last);
        diag_from(R.d0, o.d0);
      }
    }
    if (got > 0) {
      v->t = c.t; v->step = c.step; v->dt = c.dt;
      diag_from(c.diag[4], v->last);
      v->diag_valid = 0;
    }
    if (n_done) n_done[i] = got;
    if (c.nDone < c.nTarget && !(c.hasTEnd && c.t >= c.tEnd - 1e-12)) running++;
  }
  p->lastChunk = 0;
  if (n_running) *n_running = running;
  return NSL_OK;
}

// ------------------------------------------------------------------------------------------------ pool groups
// Pools of different N stepped as one: one RK4 step of every pool of the group is captured as one hipGraph (a
// fork/join capture: each pool's stream joins the origin stream's capture, enqueues its step, and rejoins), so an
// ensemble step is one graph launch whatever the mix of grids. A group launch runs each pool's chunk set-up, `steps`
// replays of the graph and each pool's record gather on the group's stream; each pool's stream then waits for it, so
// nsl_pool_ready / nsl_pool_collect of every pool work as after nsl_pool_launch. Slot arithmetic is unchanged (same
// kernels, same launch shapes), so results are bit-identical to the pools stepped alone.
struct nsl_group {
  std::vector<nsl_pool *> pools;
  hipStream_t st = 0;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

static void group_free(nsl_group *g) {
  if (!g) return;
  if (g->st) (void)hipStreamSynchronize(g->st);
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->fork) (void)hipEventDestroy(g->fork);
  if (g->done) (void)hipEventDestroy(g->done);
  for (hipEvent_t e : g->join) (void)hipEventDestroy(e);
  if (g->st) (void)hipStreamDestroy(g->st);
  delete g;
}

static int group_create_impl(nsl_pool *const *pools, int n, nsl_group **out) {
  if (!pools || !out || n < 1) return fail(NSL_ERR_ARG, "a group needs at least one pool");
  *out = nullptr;
  std::set<nsl_pool *> seen;
  for (int i = 0; i < n; i++) {
    if (!pools[i]) return fail(NSL_ERR_ARG, "null pool");
    if (!seen.insert(pools[i]).second) return fail(NSL_ERR_ARG, "a pool appears twice in the group");
    if (pools[i]->lastChunk > 0) return fail(NSL_ERR_STATE, "a pool of the group has a chunk in flight");
  }
  nsl_group *g = new nsl_group();
  g->pools.assign(pools, pools + n);
  auto bail = [&](int rc) { group_free(g); return rc; };
  if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess) return bail(fail(NSL_ERR_HIP, "hipStreamCreate failed"));
  if (hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g->done, hipEventDisableTiming) != hipSuccess)
    return bail(fail(NSL_ERR_HIP, "hipEventCreate failed"));
  g->join.resize(n, nullptr);
  for (int i = 0; i < n; i++)
    if (hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming) != hipSuccess) return bail(fail(NSL_ERR_HIP, "hipEventCreate failed"));
  for (nsl_pool *p : g->pools)
    if (hipStreamSynchronize(p->st) != hipSuccess) return bail(fail(NSL_ERR_HIP, "hipStreamSynchronize failed"));
  if (hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return bail(fail(NSL_ERR_HIP, "hipStreamBeginCapture failed"));
  int rc = hipEventRecord(g->fork, g->st) == hipSuccess ? NSL_OK : fail(NSL_ERR_HIP, "hipEventRecord failed");
  for (int i = 0; i < n && rc == NSL_OK; i++) {
    nsl_pool *p = g->pools[i];
    if (hipStreamWaitEvent(p->st, g->fork, 0) != hipSuccess) { rc = fail(NSL_ERR_HIP, "hipStreamWaitEvent failed"); break; }
    rc = step_enqueue(p->lead);
    if (rc == NSL_OK && (hipEventRecord(g->join[i], p->st) != hipSuccess || hipStreamWaitEvent(g->st, g->join[i], 0) != hipSuccess))
      rc = fail(NSL_ERR_HIP, "capture join failed");
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(g->st, &graph);
  if (rc == NSL_OK && (ec != hipSuccess || graph == nullptr)) rc = fail(NSL_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  if (rc == NSL_OK && hipGraphInstantiate(&g->gexec, graph, nullptr, nullptr, 0) != hipSuccess) rc = fail(NSL_ERR_HIP, "hipGraphInstantiate failed");
  g->graph = graph;
  (void)hipGetLastError();
  if (rc != NSL_OK) return bail(rc);
  *out = g;
  return NSL_OK;
}

static int group_launch_impl(nsl_group *g, int steps) {
  if (!g) return fail(NSL_ERR_ARG, "null group");
  if (steps < 1 || steps > NSL_RECCAP) return fail(NSL_ERR_ARG, "steps per launch must be in 1..64");
  for (nsl_pool *p : g->pools)
    if (p->lastChunk > 0) return fail(NSL_ERR_STATE, "a pool of the group has a chunk in flight: collect it first");
  // the group's stream starts after whatever each pool's stream holds (admissions, state uploads)
  for (nsl_pool *p : g->pools) {
    HIP_TRY(hipEventRecord(p->lead->evfence, p->st));
    HIP_TRY(hipStreamWaitEvent(g->st, p->lead->evfence, 0));
  }
  for (nsl_pool *p : g->pools) {
    hipLaunchKernelGGL(k_chunk_begin, dim3(1, 1, p->C), dim3(1), 0, g->st, p->lead->ctl, p->L.bytes);
    LAUNCH_CHECK();
  }
  for (int i = 0; i < steps; i++) HIP_TRY(hipGraphLaunch(g->gexec, g->st));
  for (nsl_pool *p : g->pools) {
    const int nfin = pending_rows(p->lead);   // every slot's last step's end
    if (nfin > 0) {
      hipLaunchKernelGGL(k_finalize, dim3(1, 1, p->C), dim3(NSL_NT), 0, g->st, (const double *)p->lead->partials, nfin, 4,
                         p->lead->recs, p->lead->ctl, p->L.bytes, 1);
      LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_pool_gather, dim3(p->C), dim3(256), 0, g->st, (const char *)p->base, p->L.bytes, p->L.ctl,
                       p->L.recs, steps, p->dstage);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(p->hstage, p->dstage, (size_t)p->C * sizeof(PoolStage), hipMemcpyDeviceToHost, g->st));
  }
  HIP_TRY(hipEventRecord(g->done, g->st));
  for (nsl_pool *p : g->pools) {
    HIP_TRY(hipStreamWaitEvent(p->st, g->done, 0));
    p->lastChunk = steps;
    p->launch_gen = p->gen;
  }
  return NSL_OK;
}

// ------------------------------------------------------------------------------------------------ state I/O
static int get_state_impl(nsl_solver *s, double *const planes[6]) {
  const Geo &g = s->x.g;
  const int N = g.N, NH = g.NH;
  const size_t NS = (size_t)N * N * NH;
  std::vector<double2> h((size_t)3 * g.D);
  HIP_TRY(hipMemcpyAsync(h.data(), s->S, h.size() * sizeof(double2), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  for (int p = 0; p < 6; p++) memset(planes[p], 0, NS * sizeof(double));
  for (int cz = 0; cz < g.by; cz++) {
    const int l = full_of_h(cz, g);
    for (int cy = 0; cy < g.by; cy++) {
      const int j = full_of_h(cy, g);
      const size_t mrow = ((size_t)cz * g.by + cy) * g.hp, frow = ((size_t)l * N + j) * NH;
      for (int c = 0; c < 3; c++)
        for (int i = 0; i < g.hx; i++) {
          const double2 v = h[(size_t)c * g.D + mrow + i];
          planes[2 * c][frow + i] = v.x;
          planes[2 * c + 1][frow + i] = v.y;
        }
    }
  }
  return NSL_OK;
}

static int set_state_impl(nsl_solver *s, const double *const planes[6]) {
  const Geo &g = s->x.g;
  const int N = g.N, NH = g.NH;
  std::vector<double2> h((size_t)3 * g.D);
  for (int cz = 0; cz < g.by; cz++) {
    const int l = full_of_h(cz, g);
    for (int cy = 0; cy < g.by; cy++) {
      const int j = full_of_h(cy, g);
      const size_t mrow = ((size_t)cz * g.by + cy) * g.hp, frow = ((size_t)l * N + j) * NH;
      for (int c = 0; c < 3; c++)
        for (int i = 0; i < g.hx; i++)
          h[(size_t)c * g.D + mrow + i] = make_double2(planes[2 * c][frow + i], planes[2 * c + 1][frow + i]);
    }
  }
  HIP_TRY(hipMemcpyAsync(s->S, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice, s->st));
  if (s->fp32) TRY(aux_narrow(s));
  s->diag_valid = 0;
  return rhs_current(s);
}

// ------------------------------------------------------------------------------------------------ standalone transforms
struct TmpBuf {
  void *p = nullptr;
  ~TmpBuf() { if (p) (void)hipFree(p); }
};

static int fft1_impl(int N, int sign, int64_t lines, double *re, double *im) {
  Xform x;
  TRY(xform_init(x, geo_full(N), 0));
  const size_t bytes = (size_t)lines * N * sizeof(double);
  TmpBuf dre, dim_;
  int r = NSL_OK;
  if (hipMalloc(&dre.p, bytes) != hipSuccess || hipMalloc(&dim_.p, bytes) != hipSuccess) r = fail(NSL_ERR_HIP, "hipMalloc failed");
  if (r == NSL_OK && (hipMemcpy(dre.p, re, bytes, hipMemcpyHostToDevice) != hipSuccess ||
                      hipMemcpy(dim_.p, im, bytes, hipMemcpyHostToDevice) != hipSuccess)) r = fail(NSL_ERR_HIP, "hipMemcpy failed");
  if (r == NSL_OK) {
    const int TN = N <= 4096 ? 4096 : 8192;
    int T = std::max(1, std::min(16, TN / N));
    const int lds = T * N * (int)sizeof(double2);
    const unsigned grid = (unsigned)((lines + T - 1) / T);
    if (TN == 4096) { hipLaunchKernelGGL(k_fft_lines<4096>, dim3(grid), dim3(NSL_NT), lds, 0, (double *)dre.p, (double *)dim_.p, (long long)lines, x.hp.P, T, (double)sign); }
    else { allow_lds(k_fft_lines<8192>, lds); hipLaunchKernelGGL(k_fft_lines<8192>, dim3(grid), dim3(NSL_NT), lds, 0, (double *)dre.p, (double *)dim_.p, (long long)lines, x.hp.P, T, (double)sign); }
    if (hipGetLastError() != hipSuccess) r = fail(NSL_ERR_HIP, "kernel launch failed");
  }
  if (r == NSL_OK && (hipMemcpy(re, dre.p, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                      hipMemcpy(im, dim_.p, bytes, hipMemcpyDeviceToHost) != hipSuccess)) r = fail(NSL_ERR_HIP, "hipMemcpy failed");
  xform_free(x);
  return r;
}

static int spectral_impl(int N, int forward, const double *f, const double *re, const double *im, double *fo,
                         double *reo, double *imo) {
  std::vector<int> fac;
  if (factorize_host(N, fac) < 0) return fail(NSL_ERR_ARG, bad_n_msg(N));
  if (N % 2) return fail(NSL_ERR_ARG, "grid size must be even for the real-to-complex transforms");
  Xform x;
  TRY(xform_init(x, geo_full(N), 0));
  const size_t N3 = (size_t)N * N * N, NS = (size_t)N * N * (N / 2 + 1);
  TmpBuf phys, A, B;
  int r = NSL_OK;
  if (hipMalloc(&phys.p, N3 * sizeof(double)) != hipSuccess || hipMalloc(&A.p, NS * sizeof(double2)) != hipSuccess ||
      hipMalloc(&B.p, NS * sizeof(double2)) != hipSuccess)
    r = fail(NSL_ERR_HIP, "hipMalloc failed");
  std::vector<double2> h(NS);
  if (r == NSL_OK && forward) {
    if (hipMemcpy(phys.p, f, N3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) r = fail(NSL_ERR_HIP, "hipMemcpy failed");
    if (r == NSL_OK) r = xfwd(x, (double *)phys.p, (double2 *)A.p, 1);
    if (r == NSL_OK) r = yfwd(x, (double2 *)A.p, (double2 *)B.p, 1, nullptr);
    if (r == NSL_OK) r = zfwd(x, (double2 *)B.p, (double2 *)A.p, 1, nullptr);
    if (r == NSL_OK && hipMemcpy(h.data(), A.p, NS * sizeof(double2), hipMemcpyDeviceToHost) != hipSuccess) r = fail(NSL_ERR_HIP, "hipMemcpy failed");
    if (r == NSL_OK) for (size_t i = 0; i < NS; i++) { reo[i] = h[i].x; imo[i] = h[i].y; }
  } else if (r == NSL_OK) {
    for (size_t i = 0; i < NS; i++) h[i] = make_double2(re[i], im[i]);
    if (hipMemcpy(A.p, h.data(), NS * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) r = fail(NSL_ERR_HIP, "hipMemcpy failed");
    if (r == NSL_OK) r = zinv(x, (double2 *)A.p, (double2 *)B.p, 0, 1, 0, nullptr);
    if (r == NSL_OK) r = yinv(x, (double2 *)B.p, (double2 *)A.p, 1, nullptr);
    if (r == NSL_OK) r = xinv(x, (double2 *)A.p, (double *)phys.p, 1);
    if (r == NSL_OK && hipMemcpy(fo, phys.p, N3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) r = fail(NSL_ERR_HIP, "hipMemcpy failed");
  }
  xform_free(x);
  return r;
}

#include "nsl_diag.hpp"

// ------------------------------------------------------------------------------------------------ bandwidth probe
// Plain device copy, four 16-byte elements per thread per trip (loads issued before the stores): the practical HBM
// ceiling the benchmark reports beside the 8 TB/s peak.
// One 16-byte element per thread, no loop: on gfx950 this plain form copies at 6.2-6.3 TB/s (bytes read + written)
// where grid-stride loops with 4 or 8 loads in flight per thread, or streams far apart, reach 4.2-5.7.
__global__ __launch_bounds__(256) void k_copy_probe(const double2 *__restrict__ a, double2 *__restrict__ b, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

static int copy_probe_impl(int64_t bytes, int reps, double *gbs) {
  const long long n = (long long)(bytes / (int64_t)sizeof(double2));
  if (n < 256 || reps < 1) return fail(NSL_ERR_ARG, "copy probe needs at least 4 KiB and one repetition");
  if ((n + 255) / 256 > 0x7fffffffLL) return fail(NSL_ERR_ARG, "copy probe size exceeds one launch");
  TmpBuf a, b;
  HIP_TRY(hipMalloc(&a.p, (size_t)n * sizeof(double2)));
  HIP_TRY(hipMalloc(&b.p, (size_t)n * sizeof(double2)));
  HIP_TRY(hipMemset(a.p, 0, (size_t)n * sizeof(double2)));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, 0, (const double2 *)a.p, (double2 *)b.p, n);   // warm-up
  HIP_TRY(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, 0, (const double2 *)a.p, (double2 *)b.p, n);
  HIP_TRY(hipEventRecord(e1, 0));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  LAUNCH_CHECK();
  *gbs = 2.0 * (double)n * sizeof(double2) * reps / (ms * 1e-3) / 1e9;
  return NSL_OK;
}

// ------------------------------------------------------------------------------------------------ no-allocation scratch
// Device scratch of the entry points that allocate nothing (nsl_spectrum_noalloc): one array per GPU, held until the
// process ends. Host code never passes `scratch` itself (in host code the name is a host-side stand-in, not a device
// address): each device's address comes from hipGetSymbolAddress, looked up once per device with that device
// current, and only that address reaches kernels. Every call on one GPU shares the array, so calls that overlap on
// one GPU (other threads, other solvers' streams) race; the entry points synchronise before returning.
#define NSL_NOALLOC_SCRATCH_BYTES (4u << 20)
static __device__ __align__(256) double scratch[NSL_NOALLOC_SCRATCH_BYTES / sizeof(double)];

// The scratch address on device `dev`, looked up on the first call for that device with it current (the caller's
// current device is restored afterwards).
static int noalloc_scratch(int dev, double **ptr) {
  static std::mutex mu;
  static void *addr[64] = {nullptr};
  if (dev < 0 || dev >= 64) return fail(NSL_ERR_STATE, "device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != dev) HIP_TRY(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess || p == nullptr)
      return fail(NSL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    addr[dev] = p;
  }
  *ptr = (double *)addr[dev];
  return NSL_OK;
}

// nsl_spectrum with its partials in the device scratch: the same kernels and launch shape, so the same bits.
static int spectrum_noalloc_impl(nsl_solver *s, double *Ek) {
  if (s->fp32) return fail(NSL_ERR_STATE, "the allocation-free spectrum serves FP64 solves (an FP32 solve's spectrum reads an FP64 view it allocates)");
  const size_t need = spectrum_scratch_doubles(s) * sizeof(double);
  if (need > NSL_NOALLOC_SCRATCH_BYTES)
    return fail(NSL_ERR_ARG, "spectrum of N = " + std::to_string(s->N) + " needs " + std::to_string(need) +
                                 " bytes of scratch; the allocation-free entry point holds " +
                                 std::to_string(NSL_NOALLOC_SCRATCH_BYTES) + " (use nsl_spectrum)");
  double *base = nullptr;
  TRY(noalloc_scratch(s->dev, &base));
  return spectrum_run(s, base, base + (size_t)spectrum_blocks(s) * nsl_spectrum_len(s), Ek);
}

// ------------------------------------------------------------------------------------------------ C ABI
extern "C" {

const char *nsl_last_error(void) { return g_err.c_str(); }
const char *nsl_version(void) { return "nslab-hip 0.1.0 (gfx950, FP64; reference NSLab 0.1.2)"; }

int nsl_factorize(int32_t N, int32_t *f, int32_t cap) {
  std::vector<int> v;
  if (factorize_host(N, v) < 0) return fail(NSL_ERR_ARG, bad_n_msg(N));
  if (f) for (int i = 0; i < (int)v.size() && i < cap; i++) f[i] = v[i];
  return (int)v.size();
}

int nsl_create(const nsl_opts *opts, nsl_solver **out) { return create_impl(opts, out); }
int nsl_create_fp32(const nsl_opts *opts, nsl_solver **out) { return create_impl(opts, out, true); }
int nsl_precision(const nsl_solver *s) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  return s->fp32 ? 32 : 64;
}

int nsl_set_ic(nsl_solver *s, int32_t ic, const double icp[4]) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  TRY(slot_in_flight(s));
  return set_ic(s, ic, icp);
}

void nsl_destroy(nsl_solver *s) {
  if (s && s->view) return;   // a pool's slot: released with nsl_pool_release, freed with the pool
  solver_free(s);
}

int nsl_run(nsl_solver *s, int64_t n, int32_t has_tend, double tend, nsl_steprec *recs, int64_t *n_done) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  TRY(slot_in_flight(s));
  return run_impl(s, n, has_tend, tend, recs, n_done);
}

int nsl_set_clock(nsl_solver *s, double t, int64_t step, double E0) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  s->t = t; s->step = step; s->E0 = E0;
  return NSL_OK;
}

int nsl_get_clock(nsl_solver *s, double *t, int64_t *step, double *dt) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  if (t) *t = s->t;
  if (step) *step = s->step;
  if (dt) *dt = s->dt;
  return NSL_OK;
}

int nsl_last(nsl_solver *s, nsl_rhsdiag *out) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  *out = s->last;
  return NSL_OK;
}

int nsl_norms(nsl_solver *s, double out[3]) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  double *dn = (double *)((char *)s->ctl + offsetof(Ctl, nrm));
  AuxView av(s, true);
  TRY(av.rc);
  TRY(norms_dev(s, s->S, dn));
  HIP_TRY(hipMemcpyAsync(out, dn, 3 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}

int nsl_get_state(nsl_solver *s, double *const planes[6]) {
  if (!s || !planes) return fail(NSL_ERR_ARG, "null argument");
  for (int i = 0; i < 6; i++) if (!planes[i]) return fail(NSL_ERR_ARG, "null plane");
  AuxView av(s, true);
  TRY(av.rc);
  return get_state_impl(s, planes);
}

int nsl_set_state(nsl_solver *s, const double *const planes[6]) {
  if (!s || !planes) return fail(NSL_ERR_ARG, "null argument");
  for (int i = 0; i < 6; i++) if (!planes[i]) return fail(NSL_ERR_ARG, "null plane");
  TRY(slot_in_flight(s));
  AuxView av(s, false);
  TRY(av.rc);
  return set_state_impl(s, planes);
}

int64_t nsl_device_bytes(const nsl_solver *s) { return s ? s->bytes : 0; }

int nsl_profile(nsl_solver *s, int32_t on) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  TRY(pull_ctl(s));
  s->prof = on ? 1 : 0;
  for (int i = 0; i < NSL_PROF_CLASSES; i++) { s->prof_ms[i] = 0; s->prof_cnt[i] = 0; }
  return NSL_OK;
}

int nsl_profile_read(nsl_solver *s, double *ms, int64_t *count) {
  if (!s || !ms || !count) return fail(NSL_ERR_ARG, "null argument");
  TRY(pull_ctl(s));
  for (int i = 0; i < NSL_PROF_CLASSES; i++) { ms[i] = s->prof_ms[i]; count[i] = s->prof_cnt[i]; }
  return NSL_OK;
}

int nsl_fft1(int32_t N, int32_t sign, int64_t lines, double *re, double *im) {
  if (!re || !im) return fail(NSL_ERR_ARG, "null argument");
  if (sign != 1 && sign != -1) return fail(NSL_ERR_ARG, "sign must be -1 or +1");
  if (lines <= 0) return NSL_OK;
  return fft1_impl(N, sign, lines, re, im);
}

int nsl_spectral_forward(int32_t N, const double *f, double *re, double *im) {
  if (!f || !re || !im) return fail(NSL_ERR_ARG, "null argument");
  return spectral_impl(N, 1, f, nullptr, nullptr, nullptr, re, im);
}

int nsl_spectral_inverse(int32_t N, const double *re, const double *im, double *f) {
  if (!f || !re || !im) return fail(NSL_ERR_ARG, "null argument");
  return spectral_impl(N, 0, nullptr, re, im, f, nullptr, nullptr);
}

int nsl_diagnose(nsl_solver *s, int32_t do_interp, double *out) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  AuxView av(s, true);
  TRY(av.rc);
  return diagnose_impl(s, do_interp, out);
}

int nsl_get_divmax(nsl_solver *s, double *d) {
  if (!s || !d) return fail(NSL_ERR_ARG, "null argument");
  *d = s->divMax;
  return NSL_OK;
}
int nsl_set_divmax(nsl_solver *s, double d) {
  if (!s) return fail(NSL_ERR_ARG, "null argument");
  s->divMax = d;
  return NSL_OK;
}

int nsl_spectrum_len(nsl_solver *s) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  return (int)ceil(sqrt(3.0) * (s->N / 3)) + 2;
}

int nsl_spectrum(nsl_solver *s, double *Ek) {
  if (!s || !Ek) return fail(NSL_ERR_ARG, "null argument");
  AuxView av(s, true);
  TRY(av.rc);
  return spectrum_impl(s, Ek);
}

int nsl_spectrum_noalloc(nsl_solver *s, double *Ek) {
  if (!s || !Ek) return fail(NSL_ERR_ARG, "null argument");
  return spectrum_noalloc_impl(s, Ek);
}

int nsl_interp_max(nsl_solver *s, int32_t top, int32_t iters, double out[5]) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  AuxView av(s, true);
  TRY(av.rc);
  return interp_max_impl(s, top, iters, out);
}

int nsl_field(nsl_solver *s, int32_t which, double *out) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  AuxView av(s, true);
  TRY(av.rc);
  return field_impl(s, which, out);
}

int nsl_exact_error(nsl_solver *s, double out[2]) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  AuxView av(s, true);
  TRY(av.rc);
  return exact_error_impl(s, out);
}

int nsl_slice(nsl_solver *s, int32_t kind, int32_t axis, int32_t index, float *out) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  AuxView av(s, true);
  TRY(av.rc);
  return slice_impl(s, kind, axis, index, out);
}

int nsl_image_diag(nsl_solver *s, double out[4]) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  AuxView av(s, true);
  TRY(av.rc);
  return image_diag_impl(s, out);
}

int nsl_cutoff_pile_up(int32_t kc, const double *spec, int32_t nb, double *out) {
  if (!spec || !out) return fail(NSL_ERR_ARG, "null argument");
  if (kc < 1 || nb <= kc) return fail(NSL_ERR_ARG, "spectrum of " + std::to_string(nb) + " bins does not reach kc = " + std::to_string(kc));
  return pile_up(kc, spec, nb, kEps64, out);
}

int nsl_cutoff_pile_up_eps(int32_t kc, const double *spec, int32_t nb, double eps, double *out) {
  if (!spec || !out) return fail(NSL_ERR_ARG, "null argument");
  if (kc < 1 || nb <= kc) return fail(NSL_ERR_ARG, "spectrum of " + std::to_string(nb) + " bins does not reach kc = " + std::to_string(kc));
  if (!(eps > 0 && eps < 1)) return fail(NSL_ERR_ARG, "machine epsilon must lie in (0, 1)");
  return pile_up(kc, spec, nb, eps, out);
}

int nsl_grid_size(const nsl_solver *s) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  return s->N;
}

int nsl_copy_probe(int64_t bytes, int32_t reps, double *gbs) {
  if (!gbs) return fail(NSL_ERR_ARG, "null argument");
  return copy_probe_impl(bytes, reps, gbs);
}

int nsl_device_memory(int64_t *free_bytes, int64_t *total_bytes) {
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = (int64_t)f;
  if (total_bytes) *total_bytes = (int64_t)t;
  return NSL_OK;
}

int nsl_pool_create(int32_t N, int32_t slots, nsl_pool **out) { return pool_create_impl(N, slots, out); }

static bool fast_size(int N) {
  switch (N) {
#define NSL_CASE(n) case n:
    NSL_FAST_SIZES(NSL_CASE)
#undef NSL_CASE
    return true;
    default:
      return false;
  }
}
static nsl_tune to_abi(const TuneCfg &c) { return nsl_tune{c.xpass, c.lp, c.xwaves, c.zfused, c.pt, c.xs, c.zs}; }

int nsl_tune_get(int32_t N, nsl_tune *out) {
  if (!fast_size(N)) return fail(NSL_ERR_ARG, "no size-specialised kernels for N = " + std::to_string(N));
  if (out) *out = to_abi(tune_for(N));
  return NSL_OK;
}

int nsl_tune_set(int32_t N, const nsl_tune *cfg) {
  if (!fast_size(N)) return fail(NSL_ERR_ARG, "no size-specialised kernels for N = " + std::to_string(N));
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (cfg == nullptr) {
    g_tune_over.erase(N);
    return NSL_OK;
  }
  TuneCfg c;
  c.xpass = cfg->xpass; c.lp = cfg->lp; c.xwaves = cfg->xwaves; c.zfused = cfg->zfused;
  c.pt = cfg->ptile; c.xs = cfg->xsched; c.zs = cfg->zsched;
  if (!tune_valid(N, c)) return fail(NSL_ERR_ARG, "kernel forms not compiled for N = " + std::to_string(N));
  g_tune_over[N] = c;
  return NSL_OK;
}

int nsl_tune_candidates(int32_t N, nsl_tune *out, int32_t cap) {
  if (!fast_size(N)) return fail(NSL_ERR_ARG, "no size-specialised kernels for N = " + std::to_string(N));
  const std::vector<TuneCfg> v = tune_candidates(N);
  if (out) for (int i = 0; i < (int)v.size() && i < cap; i++) out[i] = to_abi(v[i]);
  return (int)v.size();
}

int nsl_xpass_plan(int32_t N, const nsl_tune *cfg, int32_t cus, int64_t cap_pairs, int64_t *workgroups) {
  if (!cfg || !workgroups) return fail(NSL_ERR_ARG, "null argument");
  TuneCfg c;
  c.xpass = cfg->xpass; c.lp = cfg->lp; c.xwaves = cfg->xwaves; c.zfused = cfg->zfused;
  c.pt = cfg->ptile; c.xs = cfg->xsched; c.zs = cfg->zsched;
  long long g = 0;
  TRY(xnl_plan_n(N, c, cus > 0 ? cus : device_cus(), cap_pairs, &g));
  *workgroups = g;
  return NSL_OK;
}

void nsl_pool_destroy(nsl_pool *p) { pool_free(p); }

int nsl_group_create(nsl_pool *const *pools, int32_t n, nsl_group **out) { return group_create_impl(pools, n, out); }
int nsl_group_launch(nsl_group *g, int32_t steps) { return group_launch_impl(g, steps); }
void nsl_group_destroy(nsl_group *g) { group_free(g); }

int64_t nsl_pool_slot_bytes(int32_t N) {
  std::vector<int> f;
  if (factorize_host(N, f) < 0) return fail(NSL_ERR_ARG, bad_n_msg(N));
  return slot_bytes_of(N);
}

int nsl_pool_info(nsl_pool *p, int32_t *N, int32_t *slots, int32_t *used, int32_t *graphed, int64_t *bytes) {
  if (!p) return fail(NSL_ERR_ARG, "null pool");
  int u = 0;
  for (int v : p->used) u += v;
  if (N) *N = p->N;
  if (slots) *slots = p->C;
  if (used) *used = u;
  if (graphed) *graphed = p->graphed;
  if (bytes) *bytes = p->bytes;
  return NSL_OK;
}

int nsl_pool_admit(nsl_pool *p, const nsl_opts *opts, int32_t *slot) {
  int s = -1;
  const int r = pool_admit_impl(p, opts, &s);
  if (slot) *slot = s;
  return r;
}

int nsl_pool_release(nsl_pool *p, int32_t slot) { return pool_release_impl(p, slot); }

int nsl_pool_slot(nsl_pool *p, int32_t slot, nsl_solver **view) {
  TRY(pool_slot_check(p, slot, true));
  if (!view) return fail(NSL_ERR_ARG, "null argument");
  *view = p->slots[slot];
  return NSL_OK;
}

int nsl_pool_set_target(nsl_pool *p, int32_t slot, int64_t n, int32_t has_tend, double tend) {
  return pool_set_target_impl(p, slot, n, has_tend, tend);
}

int nsl_pool_launch(nsl_pool *p, int32_t steps) { return pool_launch_impl(p, steps); }

int nsl_pool_ready(nsl_pool *p) {
  if (!p) return fail(NSL_ERR_ARG, "null pool");
  const hipError_t e = hipStreamQuery(p->st);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) { (void)hipGetLastError(); return 0; }
  return fail(NSL_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int nsl_pool_collect(nsl_pool *p, nsl_steprec *recs, int32_t recs_per_slot, int64_t *n_done, int32_t *n_running) {
  int nr = 0;
  const int r = pool_collect_impl(p, recs, recs_per_slot, n_done, &nr);
  if (n_running) *n_running = nr;
  return r;
}

}  // extern "C"
