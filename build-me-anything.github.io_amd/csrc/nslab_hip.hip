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
#include <string>
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

template <typename K>
static void allow_lds(K kernel, int bytes) {
  if (bytes > 65536) (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// ------------------------------------------------------------------------------------------------ transform set
// A geometry + plan + tiling + work buffers, able to run the passes on a stream.
struct Xform {
  Geo g;
  HostPlan hp;
  Tiling tl;
  hipStream_t st = 0;
  // size-specialised engine (nsl_fastk.hpp) when N has a compiled schedule
  bool fast = false;
  int fT = 0, fLP = 0;
  double2 *twN = nullptr;   // (cos, sin)(2 pi k / N), k < N
  double *pmax = nullptr;   // per-workgroup max |omega|^2, |u|^2 of the x pass
};

static int fast_dispatch_zinv(const Xform &x, const double2 *U, double2 *X1, int mode, int nout, int gc, const Ctl *ctl);
static int fast_dispatch_yinv(const Xform &x, const double2 *X1, double2 *X2, int ncomp, const Ctl *ctl);
static int fast_dispatch_yfwd(const Xform &x, const double2 *X3, double2 *X4, int ncomp, const Ctl *ctl);
static int fast_dispatch_zfwd(const Xform &x, const double2 *X4, double2 *OUT, int ncomp, const Ctl *ctl);
static int fast_dispatch_xnl(const Xform &x, double2 *X2, Ctl *ctl);
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
static int zinv(const Xform &x, const double2 *U, double2 *X1, int mode, int nout, int gc, const Ctl *ctl) {
  if (x.fast) return fast_dispatch_zinv(x, U, X1, mode, nout, gc, ctl);
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
static int yinv(const Xform &x, const double2 *X1, double2 *X2, int ncomp, const Ctl *ctl) {
  if (x.fast) return fast_dispatch_yinv(x, X1, X2, ncomp, ctl);
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
static int yfwd(const Xform &x, const double2 *X3, double2 *X4, int ncomp, const Ctl *ctl) {
  if (x.fast) return fast_dispatch_yfwd(x, X3, X4, ncomp, ctl);
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
static int zfwd(const Xform &x, const double2 *X4, double2 *OUT, int ncomp, const Ctl *ctl) {
  if (x.fast) return fast_dispatch_zfwd(x, X4, OUT, ncomp, ctl);
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
static int xnl(const Xform &x, double2 *X2, Ctl *ctl) {
  if (x.fast) return fast_dispatch_xnl(x, X2, ctl);
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
  if (!tiling_for(g.N, x.tl)) return fail(NSL_ERR_ARG, "grid size " + std::to_string(g.N) + " exceeds the LDS tiling of this build (N <= 1024)");
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
    // one (max |omega|^2, max |u|^2) slot per x-pass workgroup: at most one workgroup per x row
    HIP_TRY(hipMalloc(&x.pmax, 2 * (size_t)g.N * g.N * sizeof(double)));
  }
  return NSL_OK;
}
static void xform_free(Xform &x) {
  if (x.hp.twd) (void)hipFree(x.hp.twd);
  if (x.twN) (void)hipFree(x.twN);
  if (x.pmax) (void)hipFree(x.pmax);
  x.hp.twd = nullptr;
  x.twN = nullptr;
  x.pmax = nullptr;
}

#include "nsl_fastdisp.hpp"

// ------------------------------------------------------------------------------------------------ solver
#define NSL_RECCAP 64

struct nsl_solver {
  int N = 0, ic = 0;
  double nu = 0, cfl = 0, dtFixed = 0, icp[4] = {0, 0, 0, 0};
  Xform x;
  hipStream_t st = 0;
  double2 *S = nullptr, *T = nullptr, *ACC = nullptr, *NL = nullptr, *X1 = nullptr, *X2 = nullptr;
  double *partials = nullptr;
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

// One rhs of U with the fused stage-`stage` update; fmode selects where k_finalize books the sums.
static int rhs_enqueue(nsl_solver *s, double2 *U, int stage, int fmode, StepRecDev *rec) {
  const Xform &x = s->x;
  TRY(prof_begin(s, NSL_PROF_ZINV)); TRY(zinv(x, U, s->X1, 1, 6, 0, s->ctl)); TRY(prof_end(s));
  TRY(prof_begin(s, NSL_PROF_YINV)); TRY(yinv(x, s->X1, s->X2, 6, s->ctl)); TRY(prof_end(s));
  TRY(prof_begin(s, NSL_PROF_XNL)); TRY(xnl(x, s->X2, s->ctl)); TRY(prof_end(s));
  TRY(prof_begin(s, NSL_PROF_YFWD)); TRY(yfwd(x, s->X2, s->X1, 3, s->ctl)); TRY(prof_end(s));
  TRY(prof_begin(s, NSL_PROF_ZFWD)); TRY(zfwd(x, s->X1, s->NL, 3, s->ctl)); TRY(prof_end(s));
  dim3 g((unsigned)s->nblk), b(NSL_NT);
  TRY(prof_begin(s, stage == 1 ? NSL_PROF_UPDATE1 : (stage == 4 ? NSL_PROF_UPDATE4 : NSL_PROF_UPDATE23)));
  switch (stage) {
    case 1: hipLaunchKernelGGL(k_update<1>, g, b, 0, s->st, s->NL, U, s->S, s->ACC, s->T, x.g, s->partials, s->ctl); break;
    case 2: hipLaunchKernelGGL(k_update<2>, g, b, 0, s->st, s->NL, U, s->S, s->ACC, s->T, x.g, s->partials, s->ctl); break;
    case 3: hipLaunchKernelGGL(k_update<3>, g, b, 0, s->st, s->NL, U, s->S, s->ACC, s->T, x.g, s->partials, s->ctl); break;
    default: hipLaunchKernelGGL(k_update<4>, g, b, 0, s->st, s->NL, U, s->S, s->ACC, s->T, x.g, s->partials, s->ctl); break;
  }
  LAUNCH_CHECK();
  TRY(prof_end(s));
  TRY(prof_begin(s, NSL_PROF_FINALIZE));
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NSL_NT), 0, s->st, s->partials, s->nblk, fmode, rec, s->ctl);
  LAUNCH_CHECK();
  TRY(prof_end(s));
  return NSL_OK;
}

// One RK4 step (nslab.js:312-333). Stage 1's rhs is the previous step's last rhs (identical input), already
// booked in diag[4] with ACC and T prepared; the step ends with the rhs of the new state.
static int step_enqueue(nsl_solver *s) {
  hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(1), 0, s->st, s->ctl);
  LAUNCH_CHECK();
  TRY(rhs_enqueue(s, s->T, 2, 1, nullptr));
  TRY(rhs_enqueue(s, s->T, 3, 2, nullptr));
  TRY(rhs_enqueue(s, s->T, 4, 3, nullptr));
  TRY(rhs_enqueue(s, s->S, 1, 4, s->recs));
  return NSL_OK;
}

// rhs of the current state outside a step (finishIC, set_state, diagnose): refreshes last, ACC and T.
static int rhs_current(nsl_solver *s) {
  hipLaunchKernelGGL(k_set_active, dim3(1), dim3(1), 0, s->st, s->ctl, 1);
  LAUNCH_CHECK();
  TRY(rhs_enqueue(s, s->S, 1, 4, nullptr));
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

// Mulberry32 (nslab.js:161): a = a + 0x6D2B79F5, then a 32-bit hash of a.
static inline double mb32_next(uint32_t &a) {
  a += 0x6D2B79F5u;
  uint32_t t = (a ^ (a >> 15)) * (1u | a);
  t = (t + (t ^ (t >> 7)) * (61u | t)) ^ t;
  return (double)(t ^ (t >> 14)) / 4294967296.0;
}

static inline int full_of_h(int c, const Geo &g) { return c < g.hy ? c : c + g.N - g.by; }
static inline int compact_of_h(int f, const Geo &g) {
  return f < g.hy ? f : (f >= g.N - g.by + g.hy ? f - (g.N - g.by) : -1);
}
static inline double kw_h(int f, int N) { return (double)(f <= N / 2 ? f : f - N); }

// Random solenoidal field (nslab.js:273-286): the PRNG draws follow the reference's loop order over the masked
// modes, which is the compact storage order. Drawn on the host (a sequential stream), then projected and scaled
// to E0 on the device.
static int ic_random(nsl_solver *s) {
  const Geo &g = s->x.g;
  const int N = g.N;
  uint32_t a = (uint32_t)(int32_t)s->icp[0];
  const double k0 = s->icp[1], E0 = s->icp[2];
  std::vector<double2> h((size_t)3 * g.D, make_double2(0.0, 0.0));
  for (int cz = 0; cz < g.by; cz++) {
    const int l = full_of_h(cz, g);
    const double kz = kw_h(l, N);
    for (int cy = 0; cy < g.by; cy++) {
      const int j = full_of_h(cy, g);
      const double ky = kw_h(j, N);
      for (int i = 0; i < g.hx; i++) {
        const double kx = i;
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 == 0) continue;
        if (i == 0 && (kz < 0 || (kz == 0 && ky < 0))) continue;
        const double k = sqrt(k2), amp = sqrt(pow(k, 4) * exp(-2 * (k / k0) * (k / k0)) / (4 * M_PI * k2));
        const size_t m = ((size_t)cz * g.by + cy) * g.hp + i;
        for (int c = 0; c < 3; c++) {
          double v[2];
          for (int part = 0; part < 2; part++) {
            double ra = mb32_next(a);
            ra = std::max(ra, 1e-12);
            const double rb = mb32_next(a);
            v[part] = amp * (sqrt(-2 * log(ra)) * cos(TWO_PI * rb));
          }
          h[(size_t)c * g.D + m] = make_double2(v[0], v[1]);
        }
        if (i == 0) {
          const int czm = compact_of_h((N - l) % N, g), cym = compact_of_h((N - j) % N, g);
          const size_t mm = ((size_t)czm * g.by + cym) * g.hp;
          for (int c = 0; c < 3; c++) {
            const double2 v = h[(size_t)c * g.D + m];
            h[(size_t)c * g.D + mm] = make_double2(v.x, -v.y);
          }
        }
      }
    }
  }
  HIP_TRY(hipMemcpyAsync(s->S, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice, s->st));
  hipLaunchKernelGGL(k_project, dim3(sweep_grid(g.D)), dim3(NSL_NT), 0, s->st, s->S, g);
  LAUNCH_CHECK();
  double *nrm = (double *)((char *)s->ctl + offsetof(Ctl, nrm));   // device scratch for the three norms
  TRY(norms_dev(s, s->S, nrm));
  hipLaunchKernelGGL(k_scale_energy, dim3(sweep_grid(3 * g.D)), dim3(NSL_NT), 0, s->st, s->S, 3 * g.D, (const double *)nrm, E0);
  LAUNCH_CHECK();
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}

static int set_ic(nsl_solver *s, int ic, const double *icp) {
  if (ic < 0 || ic > 4) return fail(NSL_ERR_ARG, "unknown initial condition " + std::to_string(ic));
  s->ic = ic;
  for (int q = 0; q < 4; q++) s->icp[q] = icp ? icp[q] : 0;
  const Geo &g = s->x.g;
  const int N = g.N;
  if (ic == NSL_IC_RANDOM) {
    TRY(ic_random(s));
  } else {
    double *phys = (double *)s->X2;
    hipLaunchKernelGGL(k_ic_phys, dim3(sweep_grid((long long)N * N * N)), dim3(NSL_NT), 0, s->st, phys, N, ic == NSL_IC_TUBES ? 3 : ic,
                       s->icp[0], s->icp[1], s->icp[2], s->icp[3]);
    LAUNCH_CHECK();
    if (ic == NSL_IC_TUBES) {
      TRY(forward3_from_phys(s, s->NL));
      hipLaunchKernelGGL(k_tubes_finish, dim3(sweep_grid(g.D)), dim3(NSL_NT), 0, s->st, s->NL, s->S, g);
      LAUNCH_CHECK();
    } else {
      TRY(forward3_from_phys(s, s->S));
      hipLaunchKernelGGL(k_project, dim3(sweep_grid(g.D)), dim3(NSL_NT), 0, s->st, s->S, g);
      LAUNCH_CHECK();
    }
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
  (void)hipFree(s->S); (void)hipFree(s->T); (void)hipFree(s->ACC); (void)hipFree(s->NL); (void)hipFree(s->X1); (void)hipFree(s->X2);
  (void)hipFree(s->partials); (void)hipFree(s->ctl); (void)hipFree(s->recs);
  if (s->phys) (void)hipFree(s->phys);
  if (s->dpart) (void)hipFree(s->dpart);
  if (s->hrecs) (void)hipHostFree(s->hrecs);
  if (s->hctl) (void)hipHostFree(s->hctl);
  xform_free(s->x);
  for (auto e : s->evs) (void)hipEventDestroy(e);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

static int create_impl(const nsl_opts *o, nsl_solver **out) {
  if (!o || !out) return fail(NSL_ERR_ARG, "null argument");
  *out = nullptr;
  const int N = o->N;
  std::vector<int> f;
  if (factorize_host(N, f) < 0) return fail(NSL_ERR_ARG, bad_n_msg(N));
  if (N % 2) return fail(NSL_ERR_ARG, "grid size must be even for the real-to-complex transforms (got " + std::to_string(N) + ")");
  if (o->ic < 0 || o->ic > 4) return fail(NSL_ERR_ARG, "unknown initial condition " + std::to_string(o->ic));
  nsl_solver *s = new nsl_solver();
  s->N = N; s->nu = o->nu; s->cfl = o->cfl; s->dtFixed = o->dt;
  hipError_t e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) { delete s; return fail(NSL_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
  int r = xform_init(s->x, geo_truncated(N), s->st);
  if (r != NSL_OK) { solver_free(s); return r; }
  const Geo &g = s->x.g;
  const size_t D3 = (size_t)3 * g.D;
  const size_t X1n = (size_t)6 * N * g.by * g.hp, X2n = (size_t)6 * N * N * g.hp;
  s->nblk = (int)((g.D + (long long)NSL_NT * NSL_MPT - 1) / ((long long)NSL_NT * NSL_MPT));
  if ((r = dalloc(s, &s->S, D3)) || (r = dalloc(s, &s->T, D3)) || (r = dalloc(s, &s->ACC, D3)) ||
      (r = dalloc(s, &s->NL, D3)) || (r = dalloc(s, &s->X1, X1n)) || (r = dalloc(s, &s->X2, X2n)) ||
      (r = dalloc(s, &s->partials, (size_t)s->nblk * 8)) || (r = dalloc(s, &s->ctl, 1)) ||
      (r = dalloc(s, &s->recs, NSL_RECCAP))) {
    solver_free(s);
    return r;
  }
  // the padding slots of every row stay zero for the life of the solver
  for (double2 *p : {s->S, s->T, s->ACC, s->NL})
    if (hipMemset(p, 0, D3 * sizeof(double2)) != hipSuccess) { solver_free(s); return fail(NSL_ERR_HIP, "hipMemset failed"); }
  if (hipHostMalloc((void **)&s->hrecs, NSL_RECCAP * sizeof(StepRecDev)) != hipSuccess ||
      hipHostMalloc((void **)&s->hctl, sizeof(Ctl)) != hipSuccess) {
    solver_free(s);
    return fail(NSL_ERR_HIP, "hipHostMalloc failed");
  }
  memset(s->hctl, 0, sizeof(Ctl));
  Ctl *c = s->hctl;
  c->N = N; c->kc = N / 3; c->nu = s->nu; c->cfl = s->cfl; c->dtFixed = s->dtFixed; c->E0ref = 1.0;
  if ((r = push_ctl(s)) != NSL_OK) { solver_free(s); return r; }
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
    hipLaunchKernelGGL(k_chunk_begin, dim3(1), dim3(1), 0, s->st, s->ctl);
    LAUNCH_CHECK();
    for (int i = 0; i < chunk; i++) TRY(step_enqueue(s));
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

int nsl_set_ic(nsl_solver *s, int32_t ic, const double icp[4]) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
  return set_ic(s, ic, icp);
}

void nsl_destroy(nsl_solver *s) { solver_free(s); }

int nsl_run(nsl_solver *s, int64_t n, int32_t has_tend, double tend, nsl_steprec *recs, int64_t *n_done) {
  if (!s) return fail(NSL_ERR_ARG, "null solver");
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
  TRY(norms_dev(s, s->S, dn));
  HIP_TRY(hipMemcpyAsync(out, dn, 3 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}

int nsl_get_state(nsl_solver *s, double *const planes[6]) {
  if (!s || !planes) return fail(NSL_ERR_ARG, "null argument");
  for (int i = 0; i < 6; i++) if (!planes[i]) return fail(NSL_ERR_ARG, "null plane");
  return get_state_impl(s, planes);
}

int nsl_set_state(nsl_solver *s, const double *const planes[6]) {
  if (!s || !planes) return fail(NSL_ERR_ARG, "null argument");
  for (int i = 0; i < 6; i++) if (!planes[i]) return fail(NSL_ERR_ARG, "null plane");
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
  return spectrum_impl(s, Ek);
}

int nsl_interp_max(nsl_solver *s, int32_t top, int32_t iters, double out[5]) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  return interp_max_impl(s, top, iters, out);
}

int nsl_field(nsl_solver *s, int32_t which, double *out) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  return field_impl(s, which, out);
}

int nsl_exact_error(nsl_solver *s, double out[2]) {
  if (!s || !out) return fail(NSL_ERR_ARG, "null argument");
  return exact_error_impl(s, out);
}

}  // extern "C"
