// diagnose(), spectrum(), interpMax(), exactError() and the physical fields of the NSLab solver
// (pocket-wind-tunnel/src/nslab.js:339-594), on the GPU. Included by nslab_hip.hip after the solver definition.
//
// Physical work buffers (phys, allocated on the first diagnose): 17 fields of N^3 float64 —
//   0..2 u v w, 3..5 omega, 6..8 gradient scratch, 9..14 Sxx Syy Szz Sxy Sxz Syz, 15 Q (G2), 16 stretching (PST).
// Every sum is a fixed-order tree (per-block partials, then a fixed reduction), every max is exact.
#pragma once

namespace nsl {

enum { PH_U = 0, PH_OM = 3, PH_G = 6, PH_S = 9, PH_Q = 15, PH_PST = 16, PH_COUNT = 17 };

// Gradient accumulation for velocity component c (nslab.js:357-368).
__global__ __launch_bounds__(NSL_NT) void k_diag_acc(double *__restrict__ ph, size_t N3, int c,
                                                      double *__restrict__ partials) {
  __shared__ double scratch[2 * NSL_NT];
  const double *gx = ph + PH_G * N3, *gy = ph + (PH_G + 1) * N3, *gz = ph + (PH_G + 2) * N3;
  const double *ox = ph + PH_OM * N3, *oy = ph + (PH_OM + 1) * N3, *oz = ph + (PH_OM + 2) * N3;
  const double *oc = ph + (PH_OM + c) * N3;
  double *Sxx = ph + PH_S * N3, *Syy = ph + (PH_S + 1) * N3, *Szz = ph + (PH_S + 2) * N3;
  double *Sxy = ph + (PH_S + 3) * N3, *Sxz = ph + (PH_S + 4) * N3, *Syz = ph + (PH_S + 5) * N3;
  double *G2 = ph + PH_Q * N3, *PST = ph + PH_PST * N3;
  double acc[2] = {0, 0};
  const size_t stride = (size_t)gridDim.x * NSL_NT;
  for (size_t n = (size_t)blockIdx.x * NSL_NT + threadIdx.x; n < N3; n += stride) {
    const double a = gx[n], b = gy[n], d = gz[n];
    const double g2 = c == 0 ? 0.0 : G2[n], pst = c == 0 ? 0.0 : PST[n];
    G2[n] = g2 + (a * a + b * b + d * d);
    PST[n] = pst + oc[n] * (ox[n] * a + oy[n] * b + oz[n] * d);
    const double dg = c == 0 ? a : (c == 1 ? b : d);
    acc[0] += dg * dg * dg; acc[1] += dg * dg;
    if (c == 0) { Sxx[n] = a; Sxy[n] = 0.0 + 0.5 * b; Sxz[n] = 0.0 + 0.5 * d; }
    else if (c == 1) { Syy[n] = b; Sxy[n] += 0.5 * a; Syz[n] = 0.0 + 0.5 * d; }
    else { Szz[n] = d; Sxz[n] += 0.5 * a; Syz[n] += 0.5 * b; }
  }
  block_sum<2>(acc, scratch, partials + (size_t)blockIdx.x * 2);
}

// Eigenvectors of the symmetric 3x3 strain (nslab.js:396-412), ordered lambda1 >= lambda2 >= lambda3.
__device__ void eig3_dev(double a, double b, double c, double d, double e, double f, double ev[3][3]) {
  const double p1 = d * d + e * e + f * f;
  if (p1 < 1e-30) {
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) ev[i][j] = i == j ? 1.0 : 0.0;
    return;
  }
  const double q = (a + b + c) / 3, p2 = (a - q) * (a - q) + (b - q) * (b - q) + (c - q) * (c - q) + 2 * p1, p = sqrt(p2 / 6);
  const double A = (a - q) / p, B = (b - q) / p, C = (c - q) / p, D = d / p, Ee = e / p, F = f / p;
  double r = 0.5 * (A * (B * C - F * F) - D * (D * C - F * Ee) + Ee * (D * F - B * Ee));
  r = fmax(-1.0, fmin(1.0, r));
  const double phi = acos(r) / 3;
  const double l1 = q + 2 * p * cos(phi), l3 = q + 2 * p * cos(phi + 2 * 3.141592653589793 / 3);
  double vecs[2][3];
  for (int w = 0; w < 2; w++) {
    const double lam = w == 0 ? l1 : l3;
    const double r0[3] = {a - lam, d, e}, r1[3] = {d, b - lam, f}, r2[3] = {e, f, c - lam};
    const double *pa[3] = {r0, r0, r1}, *pb[3] = {r1, r2, r2};
    double best[3] = {0, 0, 0}, bl = -1;
    for (int k = 0; k < 3; k++) {
      const double *x = pa[k], *y = pb[k];
      const double v0 = x[1] * y[2] - x[2] * y[1], v1 = x[2] * y[0] - x[0] * y[2], v2 = x[0] * y[1] - x[1] * y[0];
      const double l = v0 * v0 + v1 * v1 + v2 * v2;
      if (l > bl) { bl = l; best[0] = v0; best[1] = v1; best[2] = v2; }
    }
    double l = sqrt(bl);
    if (l == 0 || l != l) l = 1;
    vecs[w][0] = best[0] / l; vecs[w][1] = best[1] / l; vecs[w][2] = best[2] / l;
  }
  const double *e1 = vecs[0], *e3 = vecs[1];
  ev[0][0] = e1[0]; ev[0][1] = e1[1]; ev[0][2] = e1[2];
  ev[1][0] = e3[1] * e1[2] - e3[2] * e1[1]; ev[1][1] = e3[2] * e1[0] - e3[0] * e1[2]; ev[1][2] = e3[0] * e1[1] - e3[1] * e1[0];
  ev[2][0] = e3[0]; ev[2][1] = e3[1]; ev[2][2] = e3[2];
}

// Pointwise statistics (nslab.js:372-386): divergence max, <|u|^3>, Q, <omega.S.omega>, alignment sums and
// 16-bin histograms (integer counts, exact). partials[block][6] = u3, Pphys, align0..2, cnt.
__global__ __launch_bounds__(NSL_NT) void k_diag_final(double *__restrict__ ph, size_t N3,
                                                        double *__restrict__ partials,
                                                        unsigned long long *__restrict__ hist,
                                                        unsigned long long *__restrict__ divmax) {
  __shared__ double scratch[6 * NSL_NT];
  __shared__ unsigned int lh[48];
  if (threadIdx.x < 48) lh[threadIdx.x] = 0;
  __syncthreads();
  const double *u = ph, *v = ph + N3, *w = ph + 2 * N3;
  const double *ox = ph + PH_OM * N3, *oy = ph + (PH_OM + 1) * N3, *oz = ph + (PH_OM + 2) * N3;
  const double *Sxx = ph + PH_S * N3, *Syy = ph + (PH_S + 1) * N3, *Szz = ph + (PH_S + 2) * N3;
  const double *Sxy = ph + (PH_S + 3) * N3, *Sxz = ph + (PH_S + 4) * N3, *Syz = ph + (PH_S + 5) * N3;
  double *G2 = ph + PH_Q * N3;
  const double *PST = ph + PH_PST * N3;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  double dmax = 0;
  const size_t stride = (size_t)gridDim.x * NSL_NT;
  for (size_t n = (size_t)blockIdx.x * NSL_NT + threadIdx.x; n < N3; n += stride) {
    const double dv = fabs(Sxx[n] + Syy[n] + Szz[n]);
    dmax = fmax(dmax, dv);
    const double o2 = ox[n] * ox[n] + oy[n] * oy[n] + oz[n] * oz[n];
    acc[0] += pow(u[n] * u[n] + v[n] * v[n] + w[n] * w[n], 1.5);
    G2[n] = 0.25 * o2 - 0.5 * G2[n];
    acc[1] += PST[n];
    if (o2 > 1e-20) {
      double ev[3][3];
      eig3_dev(Sxx[n], Syy[n], Szz[n], Sxy[n], Sxz[n], Syz[n], ev);
      const double io = 1 / sqrt(o2);
      for (int e = 0; e < 3; e++) {
        const double cs = fabs((ox[n] * ev[e][0] + oy[n] * ev[e][1] + oz[n] * ev[e][2]) * io);
        acc[2 + e] += cs;
        const int bin = min(15, (int)floor(cs * 16));
        atomicAdd(&lh[e * 16 + bin], 1u);
      }
      acc[5] += 1.0;
    }
  }
  dmax = wave_max(dmax);
  block_sum<6>(acc, scratch, partials + (size_t)blockIdx.x * 6);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) atomicMax(divmax, (unsigned long long)__double_as_longlong(dmax));
  if (threadIdx.x < 48 && lh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)lh[threadIdx.x]);
}

// Sum of value v over nblk blocks of partials[block][nv] (one workgroup per value, fixed tree).
__global__ __launch_bounds__(NSL_NT) void k_sum_partials(const double *__restrict__ partials, int nblk, int nv,
                                                          double *__restrict__ out) {
  __shared__ double scratch[NSL_NT];
  __shared__ double tot[1];
  const int v = blockIdx.x;
  double acc[1] = {0};
  for (int b = threadIdx.x; b < nblk; b += NSL_NT) acc[0] += partials[(size_t)b * nv + v];
  block_sum<1>(acc, scratch, tot);
  if (threadIdx.x == 0) out[v] = tot[0];
}

// Shell spectrum (nslab.js:414-420): E(k) += 0.5 hw |u_k|^2 for k = round(|k|) < nb. Each wave keeps private bins
// in LDS and its 64 lanes add in lane order, so the sum order is fixed; partials[block][nb].
__global__ __launch_bounds__(NSL_NT) void k_spectrum(const double2 *__restrict__ U, const Geo g, int nb,
                                                      double *__restrict__ partials) {
  extern __shared__ double bins[];   // [4][nb]
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < 4 * nb; i += NSL_NT) bins[i] = 0;
  __syncthreads();
  double *mine = bins + wid * nb;
  const long long D = g.D;
  const long long per = (D + gridDim.x - 1) / gridDim.x;
  const long long lo = (long long)blockIdx.x * per, hi = min(D, lo + per);
  for (long long base = lo + wid * 64; base < hi; base += NSL_NT) {
    const long long m = base + lane;
    int bin = -1;
    double e = 0;
    if (m < hi) {
      double kx, ky, kz;
      const bool real_mode = mode_k(m, g, kx, ky, kz);
      const int k = (int)floor(sqrt(kx * kx + ky * ky + kz * kz) + 0.5);
      if (real_mode && k < nb) {
        const double2 a = U[m], b = U[D + m], c = U[2 * D + m];
        e = a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y + c.x * c.x + c.y * c.y;
        const double hw = (kx == 0 || 2 * kx == g.N) ? 1.0 : 2.0;
        e = 0.5 * hw * e;
        bin = k;
      }
    }
    for (int l = 0; l < 64; l++) {
      if (lane == l && bin >= 0) mine[bin] += e;
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += NSL_NT)
    partials[(size_t)blockIdx.x * nb + i] = ((bins[i] + bins[nb + i]) + bins[2 * nb + i]) + bins[3 * nb + i];
}

// Enstrophy profile along z (nslab.js:539): Ez[l] = sum over the plane of |omega|^2. Grid: N blocks.
__global__ __launch_bounds__(NSL_NT) void k_zprofile(const double *__restrict__ ph, int N, double *__restrict__ Ez) {
  __shared__ double scratch[NSL_NT];
  const size_t N3 = (size_t)N * N * N, plane = (size_t)N * N;
  const double *ox = ph + PH_OM * N3, *oy = ph + (PH_OM + 1) * N3, *oz = ph + (PH_OM + 2) * N3;
  const size_t b0 = (size_t)blockIdx.x * plane;
  double acc[1] = {0};
  for (size_t i = threadIdx.x; i < plane; i += NSL_NT) {
    const size_t n = b0 + i;
    acc[0] += ox[n] * ox[n] + oy[n] * oy[n] + oz[n] * oz[n];
  }
  block_sum<1>(acc, scratch, Ez + blockIdx.x);
}

// Exact-solution error (nslab.js:585-594). kind 1 tgv2d, 2 abc. out: max err, max ref (IEEE bits, exact).
__global__ void k_exact(const double *__restrict__ ph, int N, int kind, double A, double B, double C, double f,
                        unsigned long long *__restrict__ out) {
  const size_t N3 = (size_t)N * N * N;
  const double h = NSL_TWO_PI / N;
  double err = 0, ref = 0;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < N3; n += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(n % N), j = (int)((n / N) % N), l = (int)(n / ((size_t)N * N));
    const double x = i * h, y = j * h, z = l * h;
    const double u = ph[n], v = ph[N3 + n], w = ph[2 * N3 + n];
    if (kind == 1) {
      const double ue = sin(x) * cos(y) * f, ve = -cos(x) * sin(y) * f;
      err = fmax(err, fmax(fabs(u - ue), fmax(fabs(v - ve), fabs(w))));
      ref = fmax(ref, fabs(ue));
    } else {
      const double ue = (A * sin(z) + C * cos(y)) * f, ve = (B * sin(x) + A * cos(z)) * f, we = (C * sin(y) + B * cos(x)) * f;
      err = fmax(err, fmax(fabs(u - ue), fmax(fabs(v - ve), fabs(w - we))));
      ref = fmax(ref, fmax(fabs(ue), fmax(fabs(ve), fabs(we))));
    }
  }
  err = wave_max(err); ref = wave_max(ref);
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&out[0], (unsigned long long)__double_as_longlong(err));
    atomicMax(&out[1], (unsigned long long)__double_as_longlong(ref));
  }
}

// N x N plane of a physical quantity as float32 (slice, nslab.js:566-583). kind: NSL_SLICE_* (|omega|, |u|, kinetic
// energy density, Q, stretching, u, v, w, omega_x..z; other kinds give zeros). The plane is x = ix (axis 0), y = ix
// (axis 1) or z = ix (axis 2); out[b*N + a] walks the two remaining coordinates, the lower one fastest.
__global__ void k_slice(const double *__restrict__ ph, int N, int kind, int axis, int ix, float *__restrict__ out) {
  const size_t N3 = (size_t)N * N * N;
  const int nn = N * N;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += gridDim.x * blockDim.x) {
    const int a = q % N, b = q / N;
    const size_t n = axis == 2 ? ((size_t)ix * N + b) * N + a : axis == 1 ? ((size_t)b * N + ix) * N + a
                                                                         : ((size_t)b * N + a) * N + ix;
    double v = 0.0;
    switch (kind) {
      case 0: case 1: case 2: {   // magnitudes: of omega (kind 0) or u (1, 2)
        const double *f = ph + (kind == 0 ? PH_OM : PH_U) * N3;
        const double x = f[n], y = f[N3 + n], z = f[2 * N3 + n];
        const double sq = x * x + y * y + z * z;
        v = kind == 2 ? 0.5 * sq : sqrt(sq);
        break;
      }
      case 3: v = ph[PH_Q * N3 + n]; break;
      case 4: v = ph[PH_PST * N3 + n]; break;
      case 5: case 6: case 7: v = ph[(size_t)(PH_U + kind - 5) * N3 + n]; break;
      case 8: case 9: case 10: v = ph[(size_t)(PH_OM + kind - 8) * N3 + n]; break;
      default: break;             // the host never launches other kinds
    }
    out[q] = (float)v;
  }
}

// |omega|^2 on the grid (interpMax start nodes).
__global__ void k_o2(const double *__restrict__ ph, size_t N3, double *__restrict__ o2) {
  const double *ox = ph + PH_OM * N3, *oy = ph + (PH_OM + 1) * N3, *oz = ph + (PH_OM + 2) * N3;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < N3; n += (size_t)gridDim.x * blockDim.x)
    o2[n] = ox[n] * ox[n] + oy[n] * oy[n] + oz[n] * oz[n];
}

// Vorticity half-spectrum IW = i k x S (interpMax's interpolant, nslab.js:432), compact.
__global__ void k_curl(const double2 *__restrict__ U, double2 *__restrict__ W, const Geo g) {
  const long long D = g.D;
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < D; m += (long long)gridDim.x * blockDim.x) {
    double kx, ky, kz;
    if (!mode_k(m, g, kx, ky, kz)) continue;
    const double2 a = U[m], b = U[D + m], c = U[2 * D + m];
    double pr = ky * c.x - kz * b.x, pi = ky * c.y - kz * b.y;
    W[m] = make_double2(-pi, pr);
    pr = kz * a.x - kx * c.x; pi = kz * a.y - kx * c.y;
    W[D + m] = make_double2(-pi, pr);
    pr = kx * b.x - ky * a.x; pi = kx * b.y - ky * a.y;
    W[2 * D + m] = make_double2(-pi, pr);
  }
}

// interpMax staged contraction (nslab.js:447-491) for component c at x. Stage A: one thread per retained (kz, ky)
// row sums over kx in order; stage B: one thread per kz sums over ky in order; stage T: one thread over kz.
__global__ void k_tens_a(const double2 *__restrict__ IW, const Geo g, int c, double x0, double *__restrict__ A) {
  const int rows = g.by * g.by;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const double2 *w = IW + (size_t)c * g.D + (size_t)r * g.hp;
  double s0r = 0, s0i = 0, s1r = 0, s1i = 0, s2r = 0, s2i = 0;
  for (int i = 0; i < g.hx; i++) {
    const double kxx = i, ph = kxx * x0, cr = cos(ph), ci = sin(ph), hw = (i == 0 || 2 * i == g.N) ? 1.0 : 2.0;
    const double ar = hw * (w[i].x * cr - w[i].y * ci), ai = hw * (w[i].x * ci + w[i].y * cr);
    s0r += ar; s0i += ai;
    s1r += -kxx * ai; s1i += kxx * ar;
    s2r += -kxx * kxx * ar; s2i += -kxx * kxx * ai;
  }
  double *o = A + (size_t)r * 6;
  o[0] = s0r; o[1] = s0i; o[2] = s1r; o[3] = s1i; o[4] = s2r; o[5] = s2i;
}
__global__ void k_tens_b(const double *__restrict__ A, const Geo g, double x1, double *__restrict__ B) {
  const int cz = blockIdx.x * blockDim.x + threadIdx.x;
  if (cz >= g.by) return;
  double acc[3][3][2];
  for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) { acc[a][b][0] = 0; acc[a][b][1] = 0; }
  for (int cy = 0; cy < g.by; cy++) {
    const double ky = kwv(full_of(cy, g), g.N), ph = ky * x1, cy_ = cos(ph), sy = sin(ph);
    const double *base = A + ((size_t)cz * g.by + cy) * 6;
    for (int a = 0; a < 3; a++) {
      const double ar0 = base[2 * a], ai0 = base[2 * a + 1];
      const double er = ar0 * cy_ - ai0 * sy, ei = ar0 * sy + ai0 * cy_;
      for (int bb = 0; bb + a < 3; bb++) {
        const double f = bb == 0 ? 1 : (bb == 1 ? ky : -ky * ky);
        const bool rot = bb == 1;
        const double rr = rot ? -f * ei : f * er, ri = rot ? f * er : f * ei;
        acc[a][bb][0] += rr; acc[a][bb][1] += ri;
      }
    }
  }
  double *o = B + (size_t)cz * 18;
  for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) { o[(a * 3 + b) * 2] = acc[a][b][0]; o[(a * 3 + b) * 2 + 1] = acc[a][b][1]; }
}
__global__ void k_tens_t(const double *__restrict__ B, const Geo g, double x2, double *__restrict__ T) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t[27];
  for (int q = 0; q < 27; q++) t[q] = 0;
  for (int cz = 0; cz < g.by; cz++) {
    const double kz = kwv(full_of(cz, g), g.N), ph = kz * x2, cz_ = cos(ph), sz = sin(ph);
    for (int a = 0; a < 3; a++)
      for (int bb = 0; bb + a < 3; bb++) {
        const double br0 = B[(size_t)cz * 18 + (a * 3 + bb) * 2], bi0 = B[(size_t)cz * 18 + (a * 3 + bb) * 2 + 1];
        const double er = br0 * cz_ - bi0 * sz, ei = br0 * sz + bi0 * cz_;
        for (int cc = 0; cc + a + bb < 3; cc++) {
          const double f = cc == 0 ? 1 : (cc == 1 ? kz : -kz * kz);
          const bool rot = cc == 1;
          t[(a * 3 + bb) * 3 + cc] += rot ? -f * ei : f * er;
        }
      }
  }
  for (int q = 0; q < 27; q++) T[q] = t[q];
}

}  // namespace nsl

// ------------------------------------------------------------------------------------------------ host side
static int ensure_phys(nsl_solver *s) {
  if (s->phys) return NSL_OK;
  const size_t N3 = (size_t)s->N * s->N * s->N;
  TRY(dalloc(s, &s->phys, (size_t)PH_COUNT * N3));
  HIP_TRY(hipMemsetAsync(s->phys, 0, (size_t)PH_COUNT * N3 * sizeof(double), s->st));
  const int nbmax = 4096;
  TRY(dalloc(s, &s->dpart, (size_t)nbmax * 8 + 4096));
  return NSL_OK;
}

// u (3) and omega (3) of the current state into phys[0..5].
static int physical_u_om(nsl_solver *s) {
  const Xform &x = s->x;
  TRY(zinv(x, s->S, s->X1, 1, 6, 0, nullptr));
  TRY(yinv(x, s->X1, s->X2, 6, nullptr));
  TRY(xinv(x, s->X2, s->phys, 6));
  return NSL_OK;
}

static unsigned diag_blocks(size_t N3) {
  return (unsigned)std::min<size_t>(std::max<size_t>((N3 + NSL_NT - 1) / NSL_NT, 1), 2048);
}

// The shell spectrum's launch shape: nblk workgroups each leave nb bin partials, summed in a fixed order into nb
// totals. part and tot are scratch of spectrum_scratch_doubles(s) doubles (partials first, then totals).
static unsigned spectrum_blocks(const nsl_solver *s) {
  return (unsigned)std::min<long long>(std::max<long long>((s->x.g.D + 4095) / 4096, 1), 1024);
}
static size_t spectrum_scratch_doubles(nsl_solver *s) {
  return ((size_t)spectrum_blocks(s) + 1) * (size_t)nsl_spectrum_len(s);
}
// Every workgroup writes all nb of its partials and the sum writes all nb totals: nothing needs clearing first.
static int spectrum_run(nsl_solver *s, double *part, double *tot, double *out) {
  const int nb = nsl_spectrum_len(s);
  const unsigned nblk = spectrum_blocks(s);
  hipLaunchKernelGGL(k_spectrum, dim3(nblk), dim3(NSL_NT), 4 * nb * sizeof(double), s->st, s->S, s->x.g, nb, part);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)nb), dim3(NSL_NT), 0, s->st, (const double *)part, (int)nblk, nb, tot);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(out, tot, nb * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}

static int spectrum_dev(nsl_solver *s, double *out) {
  TmpBuf buf;
  HIP_TRY(hipMalloc(&buf.p, spectrum_scratch_doubles(s) * sizeof(double)));
  double *part = (double *)buf.p;
  return spectrum_run(s, part, part + (size_t)spectrum_blocks(s) * nsl_spectrum_len(s), out);
}

static int spectrum_impl(nsl_solver *s, double *Ek) { return spectrum_dev(s, Ek); }

// cutoffPileUp (nslab.js:557-564; nslab_gpu.py:357-372): how far the spectrum rises between k8 = round(0.8 kc) and
// the cutoff kc, as the largest E(k)/E(k8) over that band (>= 1 by construction). Undefined (returns 0) when E(k8)
// is not above the noise floor of a quadratic quantity in the working precision, 100 eps^2 times the largest E(k),
// k >= 1 (eps: machine epsilon of the arithmetic the spectrum came from); returns 1 with the value otherwise.
static constexpr double kEps64 = 2.220446049250313e-16;   // FP64 machine epsilon (nslab_gpu.py:64)
static constexpr double kEps32 = 1.1920929e-07;           // FP32 machine epsilon, as nslab_gpu.py:64 writes it
static int pile_up(int kc, const double *spec, int nb, double eps, double *out) {
  const int k8 = (int)floor(0.8 * kc + 0.5);
  const double base = spec[k8];
  const double peak = nb > 1 ? *std::max_element(spec + 1, spec + nb) : 0.0;
  if (!(base > 100 * eps * eps * std::max(peak, 0.0))) return 0;
  double rise = 1.0;                          // E(k8)/E(k8)
  for (int k = k8 + 1; k <= kc; k++) rise = std::max(rise, spec[k] / base);
  *out = rise;
  return 1;
}

// imageDiag's bookkeeping over the enstrophy profile Ez[0..N) along z (nslab.js:537-550). The profile is a ring:
// o[0] the circular centroid (angle of sum Ez[l] e^{i z_l}, in [0, 2 pi)), o[3] the number of maximal circular runs
// of planes above 1 % of the profile's maximum, o[1] the longest run's length in z (2 pi when no plane is above:
// nothing localised), o[2] the gap 2 pi - o[1] that run leaves to its own periodic image.
static void image_diag(const std::vector<double> &Ez, int N, double *o) {
  double mx = 0, cs = 0, sn = 0;
  for (int l = 0; l < N; l++) {
    const double z = l * TWO_PI / N;
    mx = std::max(mx, Ez[l]);
    cs += Ez[l] * cos(z);
    sn += Ez[l] * sin(z);
  }
  o[0] = fmod(fmod(atan2(sn, cs), TWO_PI) + TWO_PI, TWO_PI);
  const double thr = 0.01 * mx;
  auto above = [&](int l) { return Ez[((l % N) + N) % N] > thr; };
  int first_off = -1;
  for (int l = 0; l < N && first_off < 0; l++)
    if (!above(l)) first_off = l;
  int bands = 0, longest = 0;
  if (first_off < 0) {             // every plane above: one band around the whole ring
    bands = 1;
    longest = N;
  } else {                         // walk the ring once, starting just after a plane below the threshold
    int len = 0;
    for (int q = 1; q <= N; q++) {
      if (above(first_off + q)) {
        if (len == 0) bands++;
        longest = std::max(longest, ++len);
      } else {
        len = 0;
      }
    }
  }
  const double ext = std::min((bands ? longest : N) * TWO_PI / N, TWO_PI);
  o[1] = ext;
  o[2] = std::max(TWO_PI - ext, 0.0);
  o[3] = (double)bands;
}

// interpMax (nslab.js:421-536): start nodes chosen from the grid |omega|^2 (stable order, as the reference's sort),
// the interpolant and its derivatives evaluated on the GPU by the staged contraction, the safeguarded Newton
// control on the host.
struct TensWS {
  nsl_solver *s;
  double2 *IW;
  double *A, *B, *T;
};

static int tensors_dev(TensWS &w, const double x[3], int c, double T[27]) {
  nsl_solver *s = w.s;
  const Geo &g = s->x.g;
  const int rows = g.by * g.by;
  hipLaunchKernelGGL(k_tens_a, dim3((rows + 127) / 128), dim3(128), 0, s->st, w.IW, g, c, x[0], w.A);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tens_b, dim3((g.by + 63) / 64), dim3(64), 0, s->st, w.A, g, x[1], w.B);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tens_t, dim3(1), dim3(64), 0, s->st, w.B, g, x[2], w.T);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(T, w.T, 27 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}

struct FGH {
  double f, g[3], H[3][3];
};

static int fgh_dev(TensWS &w, const double x[3], FGH &r) {
  memset(&r, 0, sizeof r);
  for (int c = 0; c < 3; c++) {
    double T[27];
    TRY(tensors_dev(w, x, c, T));
    const double wv = T[0];
    const double gw[3] = {T[(1 * 3 + 0) * 3 + 0], T[(0 * 3 + 1) * 3 + 0], T[(0 * 3 + 0) * 3 + 1]};
    const double Hw[3][3] = {{T[(2 * 3 + 0) * 3 + 0], T[(1 * 3 + 1) * 3 + 0], T[(1 * 3 + 0) * 3 + 1]},
                             {T[(1 * 3 + 1) * 3 + 0], T[(0 * 3 + 2) * 3 + 0], T[(0 * 3 + 1) * 3 + 1]},
                             {T[(1 * 3 + 0) * 3 + 1], T[(0 * 3 + 1) * 3 + 1], T[(0 * 3 + 0) * 3 + 2]}};
    r.f += wv * wv;
    for (int p = 0; p < 3; p++) {
      r.g[p] += 2 * wv * gw[p];
      for (int q = 0; q < 3; q++) r.H[p][q] += 2 * (gw[p] * gw[q] + wv * Hw[p][q]);
    }
  }
  return NSL_OK;
}

static double det3h(const double m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
static bool solve3h(const double M[3][3], const double r[3], double out[3]) {
  const double d = det3h(M);
  if (fabs(d) < 1e-300) return false;
  for (int i = 0; i < 3; i++) {
    double A2[3][3];
    memcpy(A2, M, sizeof A2);
    for (int q = 0; q < 3; q++) A2[q][i] = r[q];
    out[i] = det3h(A2) / d;
  }
  return true;
}

struct Cand {
  double v;
  int i, j, l;
  long long ord;
};

static int interp_max_impl(nsl_solver *s, int top, int iters, double *o) {
  const int N = s->N;
  const size_t N3 = (size_t)N * N * N;
  TRY(ensure_phys(s));
  if (!s->diag_valid) TRY(physical_u_om(s));
  const Geo &g = s->x.g;
  TmpBuf o2d, iw, A, B, T;
  HIP_TRY(hipMalloc(&o2d.p, N3 * sizeof(double)));
  HIP_TRY(hipMalloc(&iw.p, (size_t)3 * g.D * sizeof(double2)));
  HIP_TRY(hipMalloc(&A.p, (size_t)g.by * g.by * 6 * sizeof(double)));
  HIP_TRY(hipMalloc(&B.p, (size_t)g.by * 18 * sizeof(double)));
  HIP_TRY(hipMalloc(&T.p, 27 * sizeof(double)));
  hipLaunchKernelGGL(k_o2, dim3(diag_blocks(N3)), dim3(NSL_NT), 0, s->st, s->phys, N3, (double *)o2d.p);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_curl, dim3(sweep_grid(g.D)), dim3(NSL_NT), 0, s->st, s->S, (double2 *)iw.p, g);
  LAUNCH_CHECK();
  std::vector<double> o2(N3);
  HIP_TRY(hipMemcpyAsync(o2.data(), o2d.p, N3 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  // candidate start nodes: local maxima over the 6 neighbours, and the largest grid values
  std::vector<Cand> cand, flat(N3);
  for (int l = 0; l < N; l++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const size_t n = ((size_t)l * N + j) * N + i;
        const double v = o2[n];
        auto at = [&](int ll, int jj, int ii) { return o2[((size_t)ll * N + jj) * N + ii]; };
        const bool lm = v >= at((l + 1) % N, j, i) && v >= at((l + N - 1) % N, j, i) && v >= at(l, (j + 1) % N, i) &&
                        v >= at(l, (j + N - 1) % N, i) && v >= at(l, j, (i + 1) % N) && v >= at(l, j, (i + N - 1) % N);
        if (lm) cand.push_back({v, i, j, l, (long long)cand.size()});
        flat[n] = {v, i, j, l, (long long)n};
      }
  auto desc = [](const Cand &a, const Cand &b) { return a.v > b.v || (a.v == b.v && a.ord < b.ord); };
  const size_t t1 = std::min<size_t>(cand.size(), top), t2 = std::min<size_t>(N3, top);
  std::partial_sort(cand.begin(), cand.begin() + t1, cand.end(), desc);
  std::partial_sort(flat.begin(), flat.begin() + t2, flat.end(), desc);
  std::vector<Cand> starts;
  for (size_t q = 0; q < t1 + t2; q++) {
    const Cand c = q < t1 ? cand[q] : flat[q - t1];
    bool dup = false;
    for (auto &st : starts) if (st.i == c.i && st.j == c.j && st.l == c.l) { dup = true; break; }
    if (!dup) { Cand c2 = c; c2.ord = (long long)starts.size(); starts.push_back(c2); }
  }
  std::sort(starts.begin(), starts.end(), desc);
  const double h = TWO_PI / N;
  TensWS w{s, (double2 *)iw.p, (double *)A.p, (double *)B.p, (double *)T.p};
  bool have = false;
  double bf = 0, bx[3] = {0, 0, 0}, bnode = 0;
  for (auto &st : starts) {
    const double v0 = st.v;
    if (have && sqrt(v0) < 0.85 * sqrt(bf)) continue;
    double x[3] = {st.i * h, st.j * h, st.l * h};
    FGH cur;
    TRY(fgh_dev(w, x, cur));
    for (int it = 0; it < iters; it++) {
      double stp[3];
      const double mg[3] = {-cur.g[0], -cur.g[1], -cur.g[2]};
      const bool ok = solve3h(cur.H, mg, stp);
      const double gn = sqrt(cur.g[0] * cur.g[0] + cur.g[1] * cur.g[1] + cur.g[2] * cur.g[2]);
      if (!ok || stp[0] * cur.g[0] + stp[1] * cur.g[1] + stp[2] * cur.g[2] <= 0) {
        const double den = std::max(gn, 1e-300);
        for (int p = 0; p < 3; p++) stp[p] = cur.g[p] / den * 0.25 * h;
      }
      double n = sqrt(stp[0] * stp[0] + stp[1] * stp[1] + stp[2] * stp[2]);
      if (n > h) { for (int p = 0; p < 3; p++) stp[p] = stp[p] * h / n; n = h; }
      double xt[3] = {x[0] + stp[0], x[1] + stp[1], x[2] + stp[2]};
      FGH trial;
      TRY(fgh_dev(w, xt, trial));
      while (trial.f < cur.f && n > 1e-9 * h) {
        for (int p = 0; p < 3; p++) stp[p] *= 0.5;
        n *= 0.5;
        for (int p = 0; p < 3; p++) xt[p] = x[p] + stp[p];
        TRY(fgh_dev(w, xt, trial));
      }
      if (trial.f < cur.f) break;
      for (int p = 0; p < 3; p++) x[p] = x[p] + stp[p];
      cur = trial;
      if (n < 1e-7 * h) break;
    }
    if (!have || cur.f > bf) { have = true; bf = cur.f; bx[0] = x[0]; bx[1] = x[1]; bx[2] = x[2]; bnode = v0; }
  }
  o[0] = sqrt(std::max(bf, 0.0));
  for (int p = 0; p < 3; p++) o[1 + p] = fmod(fmod(bx[p], TWO_PI) + TWO_PI, TWO_PI);
  o[4] = sqrt(std::max(bnode, 0.0));
  return NSL_OK;
}

static int diagnose_impl(nsl_solver *s, int do_interp, double *out) {
  const int N = s->N;
  const size_t N3 = (size_t)N * N * N;
  TRY(ensure_phys(s));
  TRY(rhs_current(s));   // last for the current state (nslab.js:340)
  TRY(physical_u_om(s));
  const Xform &x = s->x;
  const unsigned nb = diag_blocks(N3);
  double *dp = s->dpart;   // [nb][2] per component, then [nb][6]
  TmpBuf red;
  HIP_TRY(hipMalloc(&red.p, 64 * sizeof(double) + 64 * sizeof(unsigned long long)));
  double *sums = (double *)red.p;
  unsigned long long *hist = (unsigned long long *)((char *)red.p + 64 * sizeof(double));
  HIP_TRY(hipMemsetAsync(red.p, 0, 64 * sizeof(double) + 64 * sizeof(unsigned long long), s->st));
  for (int c = 0; c < 3; c++) {
    TRY(zinv(x, s->S, s->X1, 2, 3, c, nullptr));
    TRY(yinv(x, s->X1, s->X2, 3, nullptr));
    TRY(xinv(x, s->X2, s->phys + PH_G * N3, 3));
    hipLaunchKernelGGL(k_diag_acc, dim3(nb), dim3(NSL_NT), 0, s->st, s->phys, N3, c, dp);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(2), dim3(NSL_NT), 0, s->st, (const double *)dp, (int)nb, 2, sums + 2 * c);
    LAUNCH_CHECK();
  }
  unsigned long long *dmx = hist + 48;
  hipLaunchKernelGGL(k_diag_final, dim3(nb), dim3(NSL_NT), 0, s->st, s->phys, N3, dp, hist, dmx);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3(6), dim3(NSL_NT), 0, s->st, (const double *)dp, (int)nb, 6, sums + 6);
  LAUNCH_CHECK();
  TmpBuf ezb;
  HIP_TRY(hipMalloc(&ezb.p, (size_t)N * sizeof(double)));
  double *Ez = (double *)ezb.p;
  hipLaunchKernelGGL(k_zprofile, dim3((unsigned)N), dim3(NSL_NT), 0, s->st, s->phys, N, Ez);
  LAUNCH_CHECK();
  double hs[12];
  unsigned long long hh[49];
  std::vector<double> ez(N);
  HIP_TRY(hipMemcpyAsync(hs, sums, 12 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipMemcpyAsync(hh, hist, 49 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipMemcpyAsync(ez.data(), Ez, N * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  s->diag_valid = 1;
  const double skew3 = hs[0] + hs[2] + hs[4], skew2 = hs[1] + hs[3] + hs[5];
  const double u3 = hs[6], Pphys = hs[7], cnt = hs[11];
  double divMax;
  { long long bits = (long long)hh[48]; memcpy(&divMax, &bits, 8); }
  s->divMax = std::max(s->divMax, divMax);
  const int nbins = nsl_spectrum_len(s);
  double *spec = out + NSL_DIAG_LEN;
  TRY(spectrum_dev(s, spec));
  const nsl_rhsdiag d = s->last;
  const double nu = s->nu, kc = N / 3;
  const double epsc = std::max(d.eps, 1e-300);
  const double eta = pow(nu * nu * nu / epsc, 0.25), kmax = sqrt(3.0) * kc;
  const double lam = sqrt(10 * nu * d.E / epsc), Rel = sqrt(2 * d.E / 3) * lam / nu;
  const double cn = std::max(cnt, 1.0);
  out[0] = s->t; out[1] = (double)s->step; out[2] = d.E; out[3] = d.Z; out[4] = d.P; out[5] = d.eps; out[6] = d.omMax;
  out[7] = d.uMax; out[8] = d.Pspec; out[9] = Pphys / (double)N3; out[10] = d.Tnl;
  out[11] = skew2 > 0 ? (skew3 / 3) / pow(skew2 / 3, 1.5) : 0;
  out[12] = divMax; out[13] = eta; out[14] = kmax * eta; out[15] = Rel; out[16] = lam;
  out[17] = -(6 * sqrt(15.0) / 7) * d.Pspec / pow(std::max(2 * d.Z, 1e-300), 1.5);
  for (int e = 0; e < 3; e++) out[18 + e] = hs[8 + e] / cn;
  out[21] = pow(u3 / (double)N3, 1.0 / 3);
  { double pu; out[22] = pile_up(N / 3, spec, nbins, s->fp32 ? kEps32 : kEps64, &pu) ? pu : NAN; }
  image_diag(ez, N, out + 23);
  if (do_interp && N <= 128) TRY(interp_max_impl(s, 24, 12, out + 27));
  else for (int q = 27; q < 32; q++) out[q] = NAN;
  for (int e = 0; e < 3; e++) for (int b = 0; b < 16; b++) out[32 + e * 16 + b] = (double)hh[e * 16 + b] / cn;
  return NSL_OK;
}

static int field_impl(nsl_solver *s, int which, double *out) {
  if (which < 0 || which > 7) return fail(NSL_ERR_ARG, "field index must be 0..7");
  const size_t N3 = (size_t)s->N * s->N * s->N;
  TRY(ensure_phys(s));
  if (!s->diag_valid && which < 6) TRY(physical_u_om(s));
  const int slot = which < 6 ? which : (which == 6 ? PH_Q : PH_PST);
  HIP_TRY(hipMemcpyAsync(out, s->phys + (size_t)slot * N3, N3 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}

static int slice_impl(nsl_solver *s, int kind, int axis, int index, float *out) {
  const int N = s->N;
  if (kind < NSL_SLICE_VORT || kind > NSL_SLICE_OMZ) {   // an unknown kind: a plane of zeros, nothing launched
    memset(out, 0, (size_t)N * N * sizeof(float));
    return NSL_OK;
  }
  TRY(ensure_phys(s));
  if (!s->diag_valid && kind != NSL_SLICE_Q && kind != NSL_SLICE_STRETCH) TRY(physical_u_om(s));
  const int ix = std::max(0, std::min(N - 1, index));
  TmpBuf d;
  HIP_TRY(hipMalloc(&d.p, (size_t)N * N * sizeof(float)));
  hipLaunchKernelGGL(k_slice, dim3((unsigned)std::min(4096, (N * N + NSL_NT - 1) / NSL_NT)), dim3(NSL_NT), 0, s->st,
                     (const double *)s->phys, N, kind, axis, ix, (float *)d.p);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(out, d.p, (size_t)N * N * sizeof(float), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  return NSL_OK;
}

// imageDiag() of the current state: the vorticity's enstrophy profile along z on the GPU, then image_diag.
static int image_diag_impl(nsl_solver *s, double *o) {
  const int N = s->N;
  TRY(ensure_phys(s));
  if (!s->diag_valid) TRY(physical_u_om(s));
  TmpBuf ezb;
  HIP_TRY(hipMalloc(&ezb.p, (size_t)N * sizeof(double)));
  hipLaunchKernelGGL(k_zprofile, dim3((unsigned)N), dim3(NSL_NT), 0, s->st, s->phys, N, (double *)ezb.p);
  LAUNCH_CHECK();
  std::vector<double> ez(N);
  HIP_TRY(hipMemcpyAsync(ez.data(), ezb.p, N * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  image_diag(ez, N, o);
  return NSL_OK;
}

static int exact_error_impl(nsl_solver *s, double *o) {
  if (s->ic != NSL_IC_TGV2D && s->ic != NSL_IC_ABC) return 0;
  const int N = s->N;
  const size_t N3 = (size_t)N * N * N;
  TRY(ensure_phys(s));
  const Xform &x = s->x;
  TRY(zinv(x, s->S, s->X1, 0, 3, 0, nullptr));
  TRY(yinv(x, s->X1, s->X2, 3, nullptr));
  TRY(xinv(x, s->X2, s->phys, 3));
  s->diag_valid = 0;
  TmpBuf r;
  HIP_TRY(hipMalloc(&r.p, 2 * sizeof(unsigned long long)));
  HIP_TRY(hipMemsetAsync(r.p, 0, 2 * sizeof(unsigned long long), s->st));
  const int kind = s->ic == NSL_IC_TGV2D ? 1 : 2;
  const double f = kind == 1 ? exp(-2 * s->nu * s->t) : exp(-s->nu * s->t);
  hipLaunchKernelGGL(k_exact, dim3(diag_blocks(N3)), dim3(NSL_NT), 0, s->st, s->phys, N, kind, s->icp[0], s->icp[1], s->icp[2], f, (unsigned long long *)r.p);
  LAUNCH_CHECK();
  unsigned long long bits[2];
  HIP_TRY(hipMemcpyAsync(bits, r.p, sizeof bits, hipMemcpyDeviceToHost, s->st));
  HIP_TRY(hipStreamSynchronize(s->st));
  double err, ref;
  memcpy(&err, &bits[0], 8);
  memcpy(&ref, &bits[1], 8);
  o[0] = err;
  o[1] = err / std::max(ref, 1e-300);
  return 1;
}
