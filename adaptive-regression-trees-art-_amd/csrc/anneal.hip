// Restart-batched annealed split optimizer: one 8-wave workgroup runs 16
// starts (restarts) of one node in lockstep.
//
// Each iteration k (T_k = t0 (t1/t0)^(k/(iters-1))) fits both sides'
// weighted ridge models at every start, forms the soft-objective gradient and
// takes a fixed projected step theta <- retract(theta - step P(grad))
// (art/objectives.py:132-340, art/splitters.py:71-98; the annealing schedule
// and fixed step have no reference counterpart).  The node's rows stream
// through LDS once per pass for all 16 starts:
//   fit     per row chunk, a row-thread phase forms every start's gates pi;
//           then each wave runs v_mfma_f64_16x16x4 Grams of the non-constant
//           features for its 2 starts and both sides (weights
//           max(1 - pi, floor) and max(pi, floor), as the reference's two
//           weighted fits), with the bias and y columns on the VALU;
//   solve   trace-scaled ridge, Crout Cholesky and substitutions for 16 of
//           the 32 systems at a time, one 32-lane group per system;
//   stats   a row-thread phase forms residuals, gates and gradient terms c
//           for the wave's 2 starts; a (start, column, half) thread grid
//           forms X^T c from LDS.
// A final hard fit and hard statistics give each start's partition.
//
// Limits: d <= 16 and at most 16 non-constant features (one MFMA tile);
// problems come in blocks of 16 that share a node.

#include "art_common.h"
#include "art_hip.h"

namespace art {

namespace {

constexpr int kAR = 16;                            // starts per workgroup
constexpr int kAThreads = 512;                     // 8 waves, 2 starts each
constexpr int kARows = 64;                         // rows per staged chunk
constexpr int kADim = 16;                          // max input dimension
constexpr int kAStr = kADim + 1;                   // LDS row stride: x_0..x_{d-1}, y
constexpr int kAP = kADim + 1;                     // max basis size
constexpr int kATri = (kAP + 1) * (kAP + 2) / 2;   // packed augmented system
constexpr int kAGroups = kAThreads / 32;           // systems factored at once

__device__ inline double agate(double score) {
  return 1.0 / (1.0 + exp(-fmin(50.0, fmax(-50.0, score))));
}

struct AnnealShared {
  double theta[kAR][kAStr];
  double pdir[kAR][kAStr];
  double beta[kAR][2][kAP];
  double lam[kAR][2];
  int status[kAR][2];
  int bad[kAR];
  double G[2 * kAR][kATri];     // augmented systems, index 2 r + side
  double xs[kARows * kAStr];
  double qs[kAR][kARows];       // gates (fit), gradient terms (stats)
  double gw[2][kAR][kADim];     // X^T c halves
  double acc[kAR][4];           // loss, -sum c, count(pi >= 1/2), bad
  double hacc[kAR][4];          // SSE_L, SSE_R, n_right, bad
  double geo[kAR][4];
  double piv[kAGroups];
  double bc[kAGroups][2];
  signed char fa[kAP], fb[kAP], core[kAP];
  int nf, cbias;
};

struct AnnealCtx {
  const double* X;
  const double* y;
  int n, d, p;
  double ridge, floor;
};

__device__ inline void stage_chunk(const AnnealCtx& C, AnnealShared& S, int base, int rows) {
  // 32-lane group per row, 16 rows per sweep, 4 sweeps per 64-row chunk.
  constexpr int kSweeps = kARows / (kAThreads / 32);
  const int g = threadIdx.x >> 5, j = threadIdx.x & 31;
  const bool isx = j < C.d, isy = j == C.d;
  const double* src = isx ? C.X + ((size_t)base + g) * C.d + j : C.y + base + g;
  const size_t stride = isx ? 16 * (size_t)C.d : 16;
  const int lim = (isx || isy) ? rows : 0;
  double v[kSweeps];
#pragma unroll
  for (int i = 0; i < kSweeps; ++i) {
    v[i] = (g + 16 * i < lim) ? *src : 0.0;
    src += stride;
  }
  double* dst = S.xs + g * kAStr + j;
#pragma unroll
  for (int i = 0; i < kSweeps; ++i)
    if (g + 16 * i < lim) dst[16 * kAStr * i] = v[i];
}

// Both sides' augmented systems of every start; kind 0: soft gates at T,
// 1: hard gates.  Rows beyond the node get gate 2 (weight 0 on both sides).
__device__ void anneal_fit(const AnnealCtx& C, AnnealShared& S, int kind, double T, double floor) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, kq = lane >> 4;
  const int d = C.d, p = C.p, n = C.n, nf = S.nf;
  const bool live = col < nf;
  const int fc = live ? S.core[col] : 0;
  const int fa = live ? S.fa[fc] : 0, fb = live ? S.fb[fc] : -1;
  f64x4 acc[2][2];
  double e1[2][2], ey[2][2], sm[2][2][3];
  int badr[2] = {0, 0};
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
      acc[r][sd] = f64x4{0.0, 0.0, 0.0, 0.0};
      e1[r][sd] = ey[r][sd] = 0.0;
      sm[r][sd][0] = sm[r][sd][1] = sm[r][sd][2] = 0.0;
    }
  for (int base = 0; base < n; base += kARows) {
    const int rows = min(kARows, n - base);
    __syncthreads();
    stage_chunk(C, S, base, rows);
    __syncthreads();
    {  // row phase: lane = row, the wave's 2 starts
      const bool valid = lane < rows;
      const double* xr = S.xs + lane * kAStr;
      const double yv = valid ? xr[d] : 0.0;
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) {
        const int r = 2 * wave + rr;
        double pi = 2.0;
        if (valid) {
          double u = 0.0;
          for (int j = 0; j < d; ++j) u = fma(xr[j], S.theta[r][j], u);
          const double z = S.theta[r][d];
          if (kind == 1) {
            pi = (isfinite(u) && isfinite(z)) ? (double)(u >= z) : __builtin_nan("");
          } else {
            const double score = (u - z) / T;
            pi = isfinite(score) ? agate(score) : __builtin_nan("");
          }
          const bool ok = isfinite(pi) && pi >= 0.0 && pi <= 1.0;
          badr[rr] |= !ok;
          if (!ok) pi = 2.0;
          const double q0 = pi <= 1.0 ? fmax(1.0 - pi, floor) : 0.0;
          const double q1 = pi <= 1.0 ? fmax(pi, floor) : 0.0;
          sm[rr][0][0] += q0;
          sm[rr][0][1] += q0 * yv;
          sm[rr][0][2] += q0 * yv * yv;
          sm[rr][1][0] += q1;
          sm[rr][1][1] += q1 * yv;
          sm[rr][1][2] += q1 * yv * yv;
        }
        S.qs[r][lane] = pi;
      }
    }
    __syncthreads();
    // MFMA phase: 4 rows per step; the wave's 2 starts share the features.
#pragma unroll 1
    for (int sq = 0; sq < kARows / 4; ++sq) {
      const int row = 4 * sq + kq;
      const bool valid = row < rows;
      const double* xr = S.xs + row * kAStr;
      double v = 0.0;
      if (valid && live) {
        v = xr[fa];
        if (fb >= 0) v *= xr[fb];
      }
      const double yv = valid ? xr[d] : 0.0;
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) {
        const double pi = S.qs[2 * wave + rr][row];
        const double q0 = pi <= 1.0 ? fmax(1.0 - pi, floor) : 0.0;
        const double q1 = pi <= 1.0 ? fmax(pi, floor) : 0.0;
        acc[rr][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(q0 * v, v, acc[rr][0], 0, 0, 0);
        acc[rr][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(q1 * v, v, acc[rr][1], 0, 0, 0);
        e1[rr][0] += q0 * v;
        ey[rr][0] += q0 * yv * v;
        e1[rr][1] += q1 * v;
        ey[rr][1] += q1 * yv * v;
      }
    }
  }
  // Assemble the wave's 4 systems.
  const int cb = S.cbias;
#pragma unroll
  for (int rr = 0; rr < 2; ++rr) {
    const int b = wave_or(badr[rr]);
    if (lane == 0) S.bad[2 * wave + rr] = b;
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
      double* G = S.G[2 * (2 * wave + rr) + sd];
      // f64 16x16x4 output: lane (col, kq) holds rows kq + 4 i, column col.
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int a = kq + 4 * i;
        if (a < nf && col < nf && a <= col) {
          const int ca = S.core[a], cc = S.core[col];
          G[ca >= cc ? tri(ca, cc) : tri(cc, ca)] = acc[rr][sd][i];
        }
      }
      double a1 = e1[rr][sd], ay = ey[rr][sd];
      a1 += __shfl_xor(a1, 16, kWave);
      a1 += __shfl_xor(a1, 32, kWave);
      ay += __shfl_xor(ay, 16, kWave);
      ay += __shfl_xor(ay, 32, kWave);
      if (kq == 0 && live) {
        const int cf = S.core[col];
        G[tri(p, cf)] = ay;
        if (cb >= 0) G[cf >= cb ? tri(cf, cb) : tri(cb, cf)] = a1;
      }
      const double t0 = wave_sum(sm[rr][sd][0]), t1 = wave_sum(sm[rr][sd][1]),
                   t2 = wave_sum(sm[rr][sd][2]);
      if (lane == 0) {
        if (cb >= 0) {
          G[tri(cb, cb)] = t0;
          G[tri(p, cb)] = t1;
        }
        G[tri(p, p)] = t2;
      }
    }
  }
  __syncthreads();
}

// Crout Cholesky and substitutions of one system per 32-lane group (lane i
// owns row i); every thread calls this (block barriers).  Returns the
// group's status: 0, pivot index, -2 non-finite factor, -3 failed solve.
__device__ int group_solve(double* M, AnnealShared& S, int p, double& x) {
  const int g = threadIdx.x >> 5, i = threadIdx.x & 31;
  int st = 0;
  for (int k = 0; k < p; ++k) {
    double s = 0.0;
    if (i >= k && i < p) {
      const double* ri = M + tri(i, 0);
      const double* rk = M + tri(k, 0);
      s = ri[k];
      for (int j = 0; j < k; ++j) s -= ri[j] * rk[j];
    }
    if (i == k) S.piv[g] = s;
    __syncthreads();
    const double dkk = S.piv[g];
    if (st == 0 && !(dkk > 0.0)) st = k + 1;
    if (st == 0) {
      const double lkk = sqrt(dkk);
      if (i == k) M[tri(k, k)] = lkk;
      else if (i > k && i < p) M[tri(i, k)] = s / lkk;
    }
    __syncthreads();
  }
  int nfin = 0;
  if (st == 0 && i < p) {
    for (int j = 0; j <= i; ++j) nfin |= !isfinite(M[tri(i, j)]);
    nfin |= !isfinite(M[tri(p, i)]);
  }
  for (int o = 16; o > 0; o >>= 1) nfin |= __shfl_xor(nfin, o, 32);
  if (st == 0 && nfin) st = -2;
  double z = (st == 0 && i < p) ? M[tri(p, i)] : 0.0;
  for (int k = 0; k < p; ++k) {
    if (i == k && st == 0) {
      z = z / M[tri(k, k)];
      S.bc[g][k & 1] = z;
    }
    __syncthreads();
    const double zk = S.bc[g][k & 1];
    if (st == 0 && i > k && i < p) z -= M[tri(i, k)] * zk;
  }
  __syncthreads();
  x = z;
  for (int k = p - 1; k >= 0; --k) {
    if (i == k && st == 0) {
      x = x / M[tri(k, k)];
      S.bc[g][k & 1] = x;
    }
    __syncthreads();
    const double xk = S.bc[g][k & 1];
    if (st == 0 && i < k) x -= M[tri(k, i)] * xk;
  }
  int nx = (st == 0 && i < p && !isfinite(x)) ? 1 : 0;
  for (int o = 16; o > 0; o >>= 1) nx |= __shfl_xor(nx, o, 32);
  if (st == 0 && nx) st = -3;
  __syncthreads();
  return st;
}

// All 32 systems, 16 at a time: ridge lambda = ridge tr(G)/p, factor, solve.
__device__ void anneal_solve(const AnnealCtx& C, AnnealShared& S) {
  const int g = threadIdx.x >> 5, i = threadIdx.x & 31;
  const int p = C.p;
  for (int round = 0; round < 2 * kAR / kAGroups; ++round) {
    const int sys = round * kAGroups + g;
    const int r = sys >> 1, side = sys & 1;
    double* M = S.G[sys];
    double tr = 0.0;
    for (int k = i; k < p; k += 32) tr += M[tri(k, k)];
    for (int o = 16; o > 0; o >>= 1) tr += __shfl_xor(tr, o, 32);
    const double lam = C.ridge == 0.0 ? 0.0 : C.ridge * (tr / p);
    const bool pre = S.bad[r] == 0 && isfinite(tr) && isfinite(lam);
    __syncthreads();
    if (pre && i < p) M[tri(i, i)] += lam;
    __syncthreads();
    double x = 0.0;
    int st = group_solve(M, S, p, x);
    if (S.bad[r]) st = -1;
    else if (!pre) st = -2;
    if (i < p) S.beta[r][side][i] = st == 0 ? x : __builtin_nan("");
    if (i == 0) {
      S.status[r][side] = st;
      S.lam[r][side] = lam;
    }
    __syncthreads();
  }
}

// Residual statistics, gradient and direction of every start at T.
__device__ void anneal_stats(const AnnealCtx& C, AnnealShared& S, double T) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int d = C.d, p = C.p, n = C.n;
  const double inv = 1.0 / (n * T);
  const int gr = (tid >> 4) & 15, gj = tid & 15, gh = tid >> 8;  // start, column, row half
  double loss[2] = {0.0, 0.0}, gz[2] = {0.0, 0.0}, cnt[2] = {0.0, 0.0};
  double g = 0.0;
  for (int base = 0; base < n; base += kARows) {
    const int rows = min(kARows, n - base);
    __syncthreads();
    stage_chunk(C, S, base, rows);
    __syncthreads();
    {
      const bool valid = lane < rows;
      const double* xr = S.xs + lane * kAStr;
      double phi[kAP];
#pragma unroll
      for (int f = 0; f < kAP; ++f) {
        double v = 0.0;
        if (valid && f < p) {
          const int a = S.fa[f], b = S.fb[f];
          v = a >= 0 ? xr[a] : 1.0;
          if (b >= 0) v *= xr[b];
        }
        phi[f] = v;
      }
      const double yv = valid ? xr[d] : 0.0;
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) {
        const int r = 2 * wave + rr;
        double c = 0.0;
        if (valid) {
          double pl = 0.0, pr = 0.0;
#pragma unroll
          for (int f = 0; f < kAP; ++f)
            if (f < p) {
              pl = fma(phi[f], S.beta[r][0][f], pl);
              pr = fma(phi[f], S.beta[r][1][f], pr);
            }
          double u = 0.0;
          for (int j = 0; j < d; ++j) u = fma(xr[j], S.theta[r][j], u);
          const double pi = agate((u - S.theta[r][d]) / T);
          double l = yv - pl, rq = yv - pr;
          l *= l;
          rq *= rq;
          c = (rq - l) * pi * (1.0 - pi) * inv;
          loss[rr] += (1.0 - pi) * l + pi * rq;
          gz[rr] -= c;
          cnt[rr] += pi >= 0.5 ? 1.0 : 0.0;
        }
        S.qs[r][lane] = c;
      }
    }
    __syncthreads();
    if (gj < d)
      for (int r = gh; r < rows; r += 2) g = fma(S.qs[gr][r], S.xs[r * kAStr + gj], g);
  }
#pragma unroll
  for (int rr = 0; rr < 2; ++rr) {
    const double a = wave_sum(loss[rr]), b = wave_sum(gz[rr]), c = wave_sum(cnt[rr]);
    if (lane == 0) {
      S.acc[2 * wave + rr][0] = a;
      S.acc[2 * wave + rr][1] = b;
      S.acc[2 * wave + rr][2] = c;
    }
  }
  S.gw[gh][gr][gj] = g;
  __syncthreads();
  if (tid < kAR) {
    const int r = tid;
    const double* w = S.theta[r];
    const double gzr = S.acc[r][1];
    double gwv[kADim];
    double ww = 0.0, wg = 0.0;
#pragma unroll
    for (int j = 0; j < kADim; ++j) {
      gwv[j] = j < d ? S.gw[0][r][j] + S.gw[1][r][j] : 0.0;
      if (j < d) {
        ww += w[j] * w[j];
        wg += w[j] * gwv[j];
      }
    }
    const double proj = wg / ww;
    double pp = 0.0, wp = 0.0;
#pragma unroll
    for (int j = 0; j < kADim; ++j)
      if (j < d) {
        const double v = -gwv[j] + proj * w[j];
        S.pdir[r][j] = v;
        pp += v * v;
        wp += w[j] * v;
      }
    S.pdir[r][d] = -gzr;
    S.geo[r][0] = ww;
    S.geo[r][1] = wp;
    S.geo[r][2] = pp;
    S.acc[r][0] /= n;
    S.acc[r][3] = (S.status[r][0] != 0 || S.status[r][1] != 0 || !(ww > 0.0) ||
                   !isfinite(ww) || !isfinite(pp) || !isfinite(gzr) || !isfinite(S.acc[r][0]))
                      ? 1.0
                      : 0.0;
  }
  __syncthreads();
}

// [SSE_L, SSE_R, n_right, bad] of every start's hard partition and fit.
__device__ void anneal_hard_stats(const AnnealCtx& C, AnnealShared& S) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int d = C.d, p = C.p, n = C.n;
  double sl[2] = {0.0, 0.0}, sr[2] = {0.0, 0.0}, cnt[2] = {0.0, 0.0};
  int bad[2] = {0, 0};
  for (int base = 0; base < n; base += kARows) {
    const int rows = min(kARows, n - base);
    __syncthreads();
    stage_chunk(C, S, base, rows);
    __syncthreads();
    if (lane < rows) {
      const double* xr = S.xs + lane * kAStr;
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) {
        const int r = 2 * wave + rr;
        double u = 0.0;
        for (int j = 0; j < d; ++j) u = fma(xr[j], S.theta[r][j], u);
        const double z = S.theta[r][d];
        bad[rr] |= !(isfinite(u) && isfinite(z));
        const bool right = u >= z;
        const double* beta = S.beta[r][right ? 1 : 0];
        double pred = 0.0;
        for (int f = 0; f < p; ++f) {
          const int a = S.fa[f], b = S.fb[f];
          double v = a >= 0 ? xr[a] : 1.0;
          if (b >= 0) v *= xr[b];
          pred = fma(v, beta[f], pred);
        }
        const double e = xr[d] - pred;
        if (right) {
          sr[rr] += e * e;
          cnt[rr] += 1.0;
        } else {
          sl[rr] += e * e;
        }
      }
    }
  }
#pragma unroll
  for (int rr = 0; rr < 2; ++rr) {
    const double a = wave_sum(sl[rr]), b = wave_sum(sr[rr]), c = wave_sum(cnt[rr]);
    const int bb = wave_or(bad[rr]);
    if (lane == 0) {
      S.hacc[2 * wave + rr][0] = a;
      S.hacc[2 * wave + rr][1] = b;
      S.hacc[2 * wave + rr][2] = c;
      S.hacc[2 * wave + rr][3] = bb;
    }
  }
  __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(kAThreads) void anneal_kernel(art_split_args args) {
  extern __shared__ __align__(16) unsigned char anneal_smem[];
  AnnealShared& S = *reinterpret_cast<AnnealShared*>(anneal_smem);
  const art_split_args& a = args;
  const int tid = threadIdx.x;
  const int q0 = blockIdx.x * kAR;
  const int node = a.node[q0];
  const int n = a.n, d = a.d, p = a.basis.p;
  AnnealCtx C;
  C.X = a.X + (size_t)node * n * d;
  C.y = a.y + (size_t)node * n;
  C.n = n;
  C.d = d;
  C.p = p;
  C.ridge = a.ridge;
  C.floor = a.floor;
  if (tid < kAR * (d + 1)) {
    const int r = tid / (d + 1), j = tid - r * (d + 1);
    S.theta[r][j] = a.theta0[(size_t)(q0 + r) * (d + 1) + j];
  }
  if (tid == 0) {
    int nf = 0, cb = -1;
    for (int f = 0; f < p; ++f) {
      int fa, fb;
      basis_factors(a.basis, f, fa, fb);
      S.fa[f] = (signed char)fa;
      S.fb[f] = (signed char)fb;
      if (fa < 0) cb = f;
      else S.core[nf++] = (signed char)f;
    }
    S.nf = nf;
    S.cbias = cb;
  }
  __syncthreads();
  if (tid < kAR) {  // unit w
    double nw = 0.0;
    for (int j = 0; j < d; ++j) nw += S.theta[tid][j] * S.theta[tid][j];
    nw = sqrt(nw);
    for (int j = 0; j < d; ++j) S.theta[tid][j] /= nw;
  }
  __syncthreads();
  int err = 0;  // per start, held by thread r < kAR
  int iters_done = 0;
  const int K = a.anneal_iters;
  const int MI = a.max_iters;
  for (int k = 0; k < K; ++k) {
    const double T = a.t0 * pow(a.t1 / a.t0, (double)k / (double)max(K - 1, 1));
    anneal_fit(C, S, 0, T, C.floor);
    anneal_solve(C, S);
    anneal_stats(C, S, T);
    if (tid < kAR) {
      const int r = tid;
      if (!err) {
        if (k <= MI) a.history[(size_t)(q0 + r) * 5 * (MI + 1) + k] = S.acc[r][0];
        if (S.acc[r][3] != 0.0) {
          err = 1;
        } else {
          const double st = a.step;
          const double rr2 = sqrt(S.geo[r][0] + 2.0 * st * S.geo[r][1] + st * st * S.geo[r][2]);
          for (int j = 0; j < d; ++j) S.theta[r][j] = (S.theta[r][j] + st * S.pdir[r][j]) / rr2;
          S.theta[r][d] += st * S.pdir[r][d];
        }
      }
    }
    __syncthreads();
    ++iters_done;
  }
  // Normalize, hard refit (floor 0) and hard statistics.
  if (tid < kAR) {
    double nw = 0.0;
    for (int j = 0; j < d; ++j) nw += S.theta[tid][j] * S.theta[tid][j];
    nw = sqrt(nw);
    for (int j = 0; j < d; ++j) S.theta[tid][j] /= nw;
  }
  __syncthreads();
  const bool refit = a.hard_refit != 0;
  if (refit) {
    anneal_fit(C, S, 1, 1.0, 0.0);
    anneal_solve(C, S);
  }
  anneal_hard_stats(C, S);
  if (tid < kAR) {
    const int r = tid, q = q0 + r;
    double* hard = a.hard + (size_t)q * 8;
    const double sl = S.hacc[r][0], sr = S.hacc[r][1], cnt = S.hacc[r][2];
    hard[3] = cnt;
    if (refit) {
      hard[0] = (sl + sr) / n;
      hard[1] = cnt < n ? sl / (n - cnt) : 0.0;
      hard[2] = cnt > 0 ? sr / cnt : 0.0;
      hard[4] = S.lam[r][0];
      hard[5] = S.lam[r][1];
      hard[6] = (S.status[r][0] != 0 || S.status[r][1] != 0 || !isfinite(hard[0])) ? 1.0 : 0.0;
    } else {
      hard[0] = hard[1] = hard[2] = hard[4] = hard[5] = 0.0;
      hard[6] = S.hacc[r][3];
    }
    hard[7] = 0.0;
    int error = 0;
    if (err) error = 6;                                        // cholesky_failure
    else if (!refit && hard[6] != 0.0) error = 8;              // nonfinite_split
    else if (fmin(cnt, n - cnt) < a.required) error = 7;       // min_side_points
    else if (hard[6] != 0.0) error = 6;
    int* info = a.info + (size_t)q * 8;
    info[0] = error;
    info[1] = 0;
    info[2] = iters_done;
    info[3] = -1;
    info[4] = 0;
    info[5] = 0;
    info[6] = error != 0;
    info[7] = 1;
    double* sc = a.scalars + (size_t)q * 6;
    sc[0] = sc[1] = 0.0;
    sc[2] = a.alpha0;
    sc[3] = a.t1;
    sc[4] = S.lam[r][0];
    sc[5] = S.lam[r][1];
    for (int j = 0; j <= d; ++j) a.theta_out[(size_t)q * (d + 1) + j] = S.theta[r][j];
    for (int f = 0; f < p; ++f) {
      a.beta_out[((size_t)q * 2 + 0) * p + f] = S.beta[r][0][f];
      a.beta_out[((size_t)q * 2 + 1) * p + f] = S.beta[r][1][f];
    }
  }
}

}  // namespace art

using namespace art;

extern "C" int art_split_anneal_batched(const art_split_args* args, int problems,
                                        hipStream_t stream) {
  if (!args || problems < kAR || problems % kAR)
    return fail(ART_ERR_ARG, "art_split_anneal_batched: problems must be a positive multiple of 16");
  const int p = args->basis.p, d = args->d;
  const int nf = p - (args->basis.bias ? 1 : 0);
  if (d < 1 || d > kADim || nf < 1 || nf > 16 || p > kAP || args->n < 1)
    return fail(ART_ERR_ARG, "art_split_anneal_batched: needs 1 <= d <= 16 and at most 16 "
                             "non-constant features");
  if (args->anneal_iters < 1 || !(args->t0 > 0.0) || !(args->t1 > 0.0))
    return fail(ART_ERR_ARG, "art_split_anneal_batched: needs iterations and positive temperatures");
  if (!args->X || !args->y || !args->theta0 || !args->node || !args->theta_out ||
      !args->beta_out || !args->hard || !args->info || !args->history || !args->scalars)
    return fail(ART_ERR_ARG, "art_split_anneal_batched: null device pointer");
  const int smem = (int)sizeof(AnnealShared);
  hipError_t e = hipFuncSetAttribute((const void*)anneal_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, smem);
  if (e != hipSuccess) return hip_fail(e, "anneal_kernel shared memory attribute");
  anneal_kernel<<<problems / kAR, kAThreads, smem, stream>>>(*args);
  e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "anneal_kernel launch");
}
