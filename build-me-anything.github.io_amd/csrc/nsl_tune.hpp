// Kernel forms per grid size: the autotuner's table. Included by nslab_hip.hip after nsl_fastdisp.hpp.
//
// A solve's kernels depend on N only, so a solve runs the same arithmetic alone and in any pool. The forms of each
// N come, in order of precedence, from an override set through nsl_tune_set, from the file named by NSLAB_TUNE
// (read once; lines "N xpass lp xwaves zfused [pt xs zs]", '#' starts a comment), from the table the tuner (js/tune.js) wrote
// (nsl_tune_table.inc), and from the defaults below for sizes the table does not list.
#pragma once
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

struct TuneRow {
  int N, xpass, lp, xwaves, zfused, pt, xs, zs;
};
static const TuneRow kTuneTable[] = {
#include "nsl_tune_table.inc"
};

// Forms compiled for N (x pass: nsl_fastdisp.hpp; fused z-forward: N = 256, 512).
static bool tune_valid(int N, const TuneCfg &c) {
  if (c.zfused != 0 && c.zfused != 1) return false;
  if (c.zfused && !zu_has(N)) return false;
  if (!xp_has(N, c.xpass)) return false;
  if (c.xpass == NSL_XP_PAIR && !((c.lp == 1 || c.lp == 2 || c.lp == 4 || c.lp == 8) && xp_pair_lp_ok(N, c.lp))) return false;
  if (c.xpass == NSL_XP_STAGED && (c.xwaves < 0 || c.xwaves > 256)) return false;
  if ((c.pt != 0 && c.pt != 1) || (c.pt && alt_T(N) == 0)) return false;
  if (c.xs < 0 || c.xs > 2 || (c.xs && c.xpass != NSL_XP_STAGED)) return false;
  if ((c.xs == 1 && !xs_has(N)) || (c.xs == 2 && !xs2_has(N))) return false;
  if ((c.zs != 0 && c.zs != 1) || (c.zs && !(c.zfused && zs_has(N)))) return false;
  return true;
}

// Before any tuning (and at 1024^3, which the tuner's default sizes leave out): row pairs with LP from the LDS
// budget below 256^3, staged waves with the fused z-forward at 256^3, 512^3 and 1024^3, rows otherwise.
static TuneCfg tune_default(int N) {
  TuneCfg c;
  if (N == 256 || N == 512) {
    c.xpass = NSL_XP_STAGED; c.xwaves = N == 512 ? 32 : 16; c.zfused = 1;
  } else if (N == 1024) {   // fused z-forward: 4.61 against 4.54 steps/s (tubes-class TGV step)
    c.xpass = NSL_XP_STAGED; c.xwaves = 16; c.zfused = 1;
  } else if (N <= 384) {
    c.xpass = NSL_XP_PAIR; c.lp = fast_LP(N);
  } else {
    c.xpass = NSL_XP_ROWS;
  }
  return c;
}

static std::mutex g_tune_mu;
static std::map<int, TuneCfg> g_tune_over;   // nsl_tune_set
static std::map<int, TuneCfg> g_tune_file;   // NSLAB_TUNE
static bool g_tune_file_read = false;

static void tune_read_file_locked() {
  if (g_tune_file_read) return;
  g_tune_file_read = true;
  const char *path = std::getenv("NSLAB_TUNE");
  if (path == nullptr || *path == 0) return;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream is(line);
    int N;
    TuneCfg c;
    if (!(is >> N >> c.xpass >> c.lp >> c.xwaves >> c.zfused)) continue;
    if (!(is >> c.pt >> c.xs >> c.zs)) c.pt = c.xs = c.zs = 0;   // the five-field form: default tile and schedules
    if (tune_valid(N, c)) g_tune_file[N] = c;
  }
}

static TuneCfg tune_for(int N) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_read_file_locked();
  auto o = g_tune_over.find(N);
  if (o != g_tune_over.end()) return o->second;
  auto f = g_tune_file.find(N);
  if (f != g_tune_file.end()) return f->second;
  for (const TuneRow &r : kTuneTable) {
    TuneCfg c;
    c.xpass = r.xpass; c.lp = r.lp; c.xwaves = r.xwaves; c.zfused = r.zfused; c.pt = r.pt; c.xs = r.xs; c.zs = r.zs;
    if (r.N == N && tune_valid(N, c)) return c;
  }
  return tune_default(N);
}

// Every compiled form of N, the tuner's search space: the x-pass forms (with their tile or persistence), fused or
// two-kernel z-forward, and, crossed with those, the pencil tile width and the schedule variants where N has them.
// The default tile and schedules come first, in the order of the x-pass forms.
static std::vector<TuneCfg> tune_candidates(int N) {
  std::vector<TuneCfg> base;
  for (int z : {0, 1}) {
    if (z && !zu_has(N)) continue;
    for (int lp : {1, 2, 4, 8}) {
      TuneCfg c; c.xpass = NSL_XP_PAIR; c.lp = lp; c.zfused = z;
      if (tune_valid(N, c)) base.push_back(c);
    }
    { TuneCfg c; c.xpass = NSL_XP_ROWS; c.zfused = z; if (tune_valid(N, c)) base.push_back(c); }
    { TuneCfg c; c.xpass = NSL_XP_WAVE; c.zfused = z; if (tune_valid(N, c)) base.push_back(c); }
    { TuneCfg c; c.xpass = NSL_XP_PLANE; c.zfused = z; if (tune_valid(N, c)) base.push_back(c); }
    for (int w : {0, 8, 16, 32, 64}) {
      TuneCfg c; c.xpass = NSL_XP_STAGED; c.xwaves = w; c.zfused = z;
      if (tune_valid(N, c)) base.push_back(c);
    }
  }
  std::vector<TuneCfg> out;
  for (int pt : {0, 1})
    for (int xs : {0, 1, 2})
      for (int zs : {0, 1})
        for (TuneCfg c : base) {
          c.pt = pt; c.xs = xs; c.zs = zs;
          if (tune_valid(N, c)) out.push_back(c);
        }
  return out;
}
