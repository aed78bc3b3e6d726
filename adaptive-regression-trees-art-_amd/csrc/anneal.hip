// Restart-batched annealed split optimizer: one 8-wave workgroup runs 16
// starts (restarts) of one node in lockstep.
//
// Each iteration k (T_k = t0 (t1/t0)^(k/(iters-1))) fits both sides'
// weighted ridge models at every start, forms the soft-objective gradient and
// takes a fixed projected step theta <- retract(theta - step P(grad))
// (art/objectives.py:132-340, art/splitters.py:71-98; the annealing schedule
// and fixed step have no reference counterpart).  The node's rows stream
// through LDS in 64-row chunks, the next chunk's loads in flight while the
// current one is processed; every contraction is an FP64 MFMA
// (v_mfma_f64_16x16x4):
//   fit     U = X W (rows x starts) gives the gates pi and the side weights
//           q = max(1 - pi, floor), max(pi, floor), as the reference's two
//           weighted fits; then all 32 augmented systems at once as one
//           product G (32 start-sides x 171) = Q^T (32 x rows) Psi (rows x 171),
//           Psi_r = the packed outer product of [phi(x_r), y_r];
//   solve   trace-scaled ridge, Crout Cholesky and substitutions for 16 of
//           the 32 systems at a time, one 32-lane group per system;
//   stats   U = X W again, predictions P = Phi B (rows x 32 start-sides),
//           residual terms c per (row, start) in registers and the gradient
//           C^T X (starts x dims).
// A final hard fit and hard statistics give each start's partition.
//
// Limits: d <= 16 and p <= 17 (so [phi, y] has at most 171 packed products);
// problems come in blocks of 16 that share a node.

#include <cstddef>

#include "art_common.h"
#include "art_hip.h"

namespace art {

namespace {

constexpr int kAR = 16;                            // starts per workgroup
constexpr int kAThreads = 512;                     // 8 waves
constexpr int kARows = 64;                         // rows per staged chunk
constexpr int kADim = 16;                          // max input dimension
constexpr int kAStr = kADim + 1;                   // LDS row stride: x_0..x_{d-1}, y
constexpr int kAP = kADim + 1;                     // max basis size
constexpr int kATri = (kAP + 1) * (kAP + 2) / 2;   // packed augmented system
constexpr int kAGroups = kAThreads / 32;           // systems factored at once
constexpr int kFStr = kAP + 2;                     // F row stride (odd)
constexpr int kQStr = 2 * kAR + 1;                 // Q / P row stride (odd)
constexpr int kSlots = (kARows * (kADim + 1) + kAThreads - 1) / kAThreads;
constexpr int kQBStr = kARows + 8;                 // bf16 side-weight plane row stride

// Gate of a score s = (u - z) * (1/T): 1 / (1 + e^-s) with s clipped to
// [-50, 50], the quotient from the hardware reciprocal and two Newton steps
// (within an ulp or two of the division).
__device__ inline double fast_gate(double s) {
  const double a = 1.0 + exp(-fmin(50.0, fmax(-50.0, s)));
  double r = __builtin_amdgcn_rcp(a);
  r = fma(r, fma(-a, r, 1.0), r);
  r = fma(r, fma(-a, r, 1.0), r);
  return r;
}

__device__ inline f64x4 mfma4(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

struct AnnealShared {
  double theta[kAR][kAStr];
  double pdir[kAR][kAStr];
  double beta[kAR][2][kAP];
  double lam[kAR][2];
  int status[kAR][2];
  int bad[kAR];
  int badw[4][kAR];
  double acc[kAR][4];           // loss, -sum c, count(pi >= 1/2), bad
  double hacc[kAR][4];          // SSE_L, SSE_R, n_right, bad
  double geo[kAR][4];
  double gsum[kAR][kADim];
  double piv[kAGroups];
  double bc[kAGroups][2];
  double xs[2][kARows * kAStr];  // double-buffered row chunks
  signed char fa[kAP], fb[kAP];
  union __attribute__((aligned(16))) {
    double G[2 * kAR][kATri];    // augmented systems, index 2 r + side
    struct {
      double F[kARows][kFStr];   // [phi, y] per row
      double Q[kARows][kQStr];   // side weights, column 2 r + side
    } fit;
    struct {
      double F[kARows][kFStr];
      // Side weights split in bf16 hi, mid, lo planes, start-side major: a
      // lane's 8 rows of one start-side are 16 contiguous bytes.
      unsigned short QB[3][2 * kAR][kQBStr];
    } fitb;
    struct {
      double pi[kARows][kAR + 1];
      double P[kARows][kQStr];   // predictions, column 2 r + side
    } st;
    double red[8][kAR][kADim + 3];  // per-wave gradient tiles and sums
  } u;
};

// The bf16 side-weight planes are read 16 bytes at a time.
static_assert(offsetof(AnnealShared, u) % 16 == 0, "LDS union must be 16-byte aligned");
static_assert((kFStr * kARows * sizeof(double)) % 16 == 0 && (kQBStr * 2) % 16 == 0,
              "bf16 plane rows must start on 16-byte boundaries");

typedef const __attribute__((address_space(1))) double gf64;

struct AnnealCtx {
  gf64* X;
  gf64* y;
  int n, d, p;
  double ridge, floor;
};

// Chunk streaming: slot i of a thread holds element tid + 512 i of the
// chunk's x values (row-major, contiguous in X) followed by its y values.
// Chunks alternate between the two xs buffers; buf names the current one.
struct Stream {
  int row[kSlots], col[kSlots];  // row < 0: unused slot; col == d: y
};

__device__ inline void stream_init(const AnnealCtx& C, Stream& X) {
  const int nx = kARows * C.d;
#pragma unroll
  for (int i = 0; i < kSlots; ++i) {
    const int e = threadIdx.x + kAThreads * i;
    if (e < nx) {
      X.row[i] = e / C.d;
      X.col[i] = e - X.row[i] * C.d;
    } else if (e < nx + kARows) {
      X.row[i] = e - nx;
      X.col[i] = C.d;
    } else {
      X.row[i] = -1;
      X.col[i] = 0;
    }
  }
}

__device__ inline void stream_load(const AnnealCtx& C, const Stream& X, int base, double* v) {
  const int rows = min(kARows, C.n - base);
#pragma unroll
  for (int i = 0; i < kSlots; ++i) {
    v[i] = 0.0;
    const int r = X.row[i];
    if (r >= 0 && r < rows)
      v[i] = X.col[i] < C.d ? C.X[(size_t)(base + r) * C.d + X.col[i]] : C.y[base + r];
  }
}

__device__ inline void stream_store(const Stream& X, double* xs, const double* v) {
#pragma unroll
  for (int i = 0; i < kSlots; ++i)
    if (X.row[i] >= 0) xs[X.row[i] * kAStr + X.col[i]] = v[i];
}

// Synchronous staging of one chunk into xs (used once, by the hard stats).
__device__ inline void stage_chunk(const AnnealCtx& C, const Stream& X, double* xs, int base) {
  double v[kSlots];
  stream_load(C, X, base, v);
  stream_store(X, xs, v);
}

// U tile of the wave (rows 16 w .. 16 w + 15 of the chunk, w < 4) against
// every start's w: lane (hi, lo) gets u for rows 16 w + hi + 4 r, start lo.
__device__ inline f64x4 gate_scores(const AnnealShared& S, const double* xs, int d, int w) {
  const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
  f64x4 u = {0.0, 0.0, 0.0, 0.0};
  for (int ks = 0; ks < (d + 3) / 4; ++ks) {
    const int j = 4 * ks + hi;
    const double a = j < d ? xs[(16 * w + lo) * kAStr + j] : 0.0;
    const double b = j < d ? S.theta[lo][j] : 0.0;
    u = mfma4(a, b, u);
  }
  return u;
}

// One chunk of the wave's NT product tiles: acc[s][mt] += Q^T (rows x
// start-sides 16 mt ..) times Psi (rows x columns of tile s), Psi entries
// F[row][oa] F[row][ob].  Operands of the next 4-row step are loaded before
// the current step's MFMAs.
template <int NT>
__device__ inline void gram_chunk(const AnnealShared& S, int lo, int hi, const int* oa,
                                  const int* ob, f64x4 (&acc)[2][2]) {
  const double* q = &S.u.fit.Q[hi][0];
  const double* f = &S.u.fit.F[hi][0];
  double a0 = q[lo], a1 = q[16 + lo], fa[NT], fb[NT];
#pragma unroll
  for (int s = 0; s < NT; ++s) {
    fa[s] = f[oa[s]];
    fb[s] = f[ob[s]];
  }
#pragma unroll 4
  for (int ks = 0; ks < kARows / 4; ++ks) {
    q += 4 * kQStr;
    f += 4 * kFStr;
    const bool more = ks + 1 < kARows / 4;
    double n0 = 0.0, n1 = 0.0, na[NT], nb[NT];
    if (more) {
      n0 = q[lo];
      n1 = q[16 + lo];
    }
#pragma unroll
    for (int s = 0; s < NT; ++s) {
      na[s] = more ? f[oa[s]] : 0.0;
      nb[s] = more ? f[ob[s]] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < NT; ++s) {
      const double bv = fa[s] * fb[s];
      acc[s][0] = mfma4(a0, bv, acc[s][0]);
      acc[s][1] = mfma4(a1, bv, acc[s][1]);
    }
    a0 = n0;
    a1 = n1;
#pragma unroll
    for (int s = 0; s < NT; ++s) {
      fa[s] = na[s];
      fb[s] = nb[s];
    }
  }
}

// Both sides' augmented systems of every start; kind 0: soft gates at T,
// 1: hard gates.  Rows beyond the node and rows with an invalid gate get
// weight 0 on both sides; an invalid gate marks its start bad.
// The pass functions are not inlined so that each gets its own register
// allocation (the kernel keeps 2 workgroups of 8 waves per CU).  Both
// streaming passes take the current xs buffer and return the next pass's.
__device__ __attribute__((noinline)) int anneal_fit(AnnealCtx C, AnnealShared& S, int buf,
                                                    int kind, double T, double floor) {
  Stream X;
  stream_init(C, X);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lo = lane & 15, hi = lane >> 4;
  const int d = C.d, p = C.p, m = p + 1, n = C.n;
  const int ntri = m * (m + 1) / 2, ntiles = (ntri + 15) >> 4;
  const double invT = 1.0 / T;
  // The wave's product tiles nt = wave, wave + 8 (nw of them); lane column
  // t = 16 nt + lo is Psi entry (a, b), t = tri(a, b); padding columns read
  // the zero column kFStr - 1 of F.
  const int nw = (wave + 8 < ntiles) ? 2 : (wave < ntiles ? 1 : 0);
  int oa[2], ob[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int t = 16 * (wave + 8 * s) + lo;
    int a = kFStr - 1, b = kFStr - 1;
    if (t < ntri) {
      a = 0;
      while (tri(a + 1, 0) <= t) ++a;
      b = t - tri(a, 0);
    }
    oa[s] = a;
    ob[s] = b;
  }
  f64x4 acc[2][2];
#pragma unroll
  for (int s = 0; s < 2; ++s) acc[s][0] = acc[s][1] = f64x4{0.0, 0.0, 0.0, 0.0};
  int bad = 0;
  const int chunks = (n + kARows - 1) / kARows;
  for (int c = 0; c < chunks; ++c) {
    const int rows = min(kARows, n - c * kARows);
    const double* xs = S.xs[buf];
    double pf[kSlots];
    stream_load(C, X, (c + 1 < chunks ? c + 1 : 0) * kARows, pf);
    if (wave < 4) {
      const f64x4 u = gate_scores(S, xs, d, wave);
      const double z = S.theta[lo][d];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * wave + hi + 4 * r;
        double q0 = 0.0, q1 = 0.0;
        if (row < rows) {
          const double uu = u[r];
          double pi;
          if (kind == 1) {
            pi = (isfinite(uu) && isfinite(z)) ? (double)(uu >= z) : __builtin_nan("");
          } else {
            const double score = (uu - z) * invT;
            pi = isfinite(score) ? fast_gate(score) : __builtin_nan("");
          }
          const bool ok = isfinite(pi) && pi >= 0.0 && pi <= 1.0;
          bad |= !ok;
          if (ok) {
            q0 = fmax(1.0 - pi, floor);
            q1 = fmax(pi, floor);
          }
        }
        S.u.fit.Q[row][2 * lo] = q0;
        S.u.fit.Q[row][2 * lo + 1] = q1;
      }
    } else {
      // [phi, y, 0 ...] of every row: 4 threads per row.
      const int t = tid - 256, row = t >> 2;
      const double* xr = xs + row * kAStr;
      for (int f = t & 3; f < kFStr; f += 4) {
        double v = 0.0;
        if (f == p) {
          v = xr[d];
        } else if (f < p) {
          const int a = S.fa[f], b = S.fb[f];
          v = a >= 0 ? xr[a] : 1.0;
          if (b >= 0) v *= xr[b];
        }
        S.u.fit.F[row][f] = v;
      }
    }
    __syncthreads();
    if (nw == 2) gram_chunk<2>(S, lo, hi, oa, ob, acc);
    else if (nw == 1) gram_chunk<1>(S, lo, hi, oa, ob, acc);
    stream_store(X, S.xs[buf ^ 1], pf);
    __syncthreads();
    buf ^= 1;
  }
  // The systems overwrite F and Q: every wave passed the loop's last barrier.
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int t = 16 * (wave + 8 * s) + lo;
    if (s < nw && t < ntri)
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) S.u.G[16 * mt + hi + 4 * r][t] = acc[s][mt][r];
  }
  if (wave < 4) {
    bad |= __shfl_xor(bad, 16, kWave);
    bad |= __shfl_xor(bad, 32, kWave);
    if (hi == 0) S.badw[wave][lo] = bad;
  }
  __syncthreads();
  if (tid < kAR) S.bad[tid] = S.badw[0][tid] | S.badw[1][tid] | S.badw[2][tid] | S.badw[3][tid];
  __syncthreads();
  return buf;
}

// One chunk of the wave's NT product tiles in split bf16: per 32-row step,
// the side-weight planes of both start-side halves are read from LDS once;
// then, tile by tile, the lane's 8 rows of psi (column 16 tile + lo) are
// formed and split in registers and the 6 products of the tile's two output
// tiles run interleaved (two independent accumulator chains) before being
// added into FP64.  One tile at a time keeps the pass within 128 VGPRs.
template <int NT>
__device__ inline void bf16_chunk(const AnnealShared& S, int lo, int hi, const int* oa,
                                  const int* ob, const NegPairs& neg, f64x4 (&acc)[2][2]) {
  constexpr int ta[6] = {0, 2, 1, 0, 1, 0}, tb[6] = {2, 0, 1, 1, 0, 0};
#pragma unroll
  for (int ks = 0; ks < kARows / 32; ++ks) {
    const int r0 = 32 * ks + 8 * hi;  // this lane's 8 rows of the step
    bf16x8 A[2][3];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int pl = 0; pl < 3; ++pl)
        A[mt][pl] = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const u32x4*>(&S.u.fitb.QB[pl][16 * mt + lo][r0]));
#pragma unroll
    for (int s = 0; s < NT; ++s) {
      u32x4 bw[3];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double* f0 = S.u.fitb.F[r0 + 2 * k];
        const double* f1 = S.u.fitb.F[r0 + 2 * k + 1];
        unsigned h, md, l;
        split3_pair((float)(f0[oa[s]] * f0[ob[s]]), (float)(f1[oa[s]] * f1[ob[s]]), neg, h,
                    md, l);
        bw[0][k] = h;
        bw[1][k] = md;
        bw[2][k] = l;
      }
      bf16x8 B[3];
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) B[pl] = __builtin_bit_cast(bf16x8, bw[pl]);
      // Products small first: hi*lo, lo*hi, mid*mid, hi*mid, mid*hi, hi*hi.
      f32x4 z[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int term = 0; term < 6; ++term)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
          z[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mt][ta[term]], B[tb[term]], z[mt],
                                                         0, 0, 0);
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[s][mt][r] += (double)z[mt][r];
    }
  }
}

// The soft fit with split-bf16 child systems (child_precision MIXED):
// the same gates, side weights and [phi, y] rows as anneal_fit, but the 32
// systems Q^T Psi come from v_mfma_f32_16x16x32_bf16 over 32-row steps.
// Each side weight q (split once, as it is written) and each packed product
// psi (split in the lane that uses it) is hi + mid + lo in bf16 (exact to
// 2^-27 of its FP32 value); the six products hi*lo, lo*hi, mid*mid, hi*mid,
// mid*hi, hi*hi accumulate in FP32 over a step and are added into FP64.
// The output tile holds start-side 16 mt + 4 hi + r, column 16 tile + lo.
__device__ __attribute__((noinline)) int anneal_fit_bf16(AnnealCtx C, AnnealShared& S, int buf,
                                                         double T, double floor) {
  Stream X;
  stream_init(C, X);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lo = lane & 15, hi = lane >> 4;
  const int d = C.d, p = C.p, m = p + 1, n = C.n;
  const int ntri = m * (m + 1) / 2, ntiles = (ntri + 15) >> 4;
  const double invT = 1.0 / T;
  const int nw = (wave + 8 < ntiles) ? 2 : (wave < ntiles ? 1 : 0);
  int oa[2], ob[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int t = 16 * (wave + 8 * s) + lo;
    int a = kFStr - 1, b = kFStr - 1;
    if (t < ntri) {
      a = 0;
      while (tri(a + 1, 0) <= t) ++a;
      b = t - tri(a, 0);
    }
    oa[s] = a;
    ob[s] = b;
  }
  const NegPairs neg = neg_pairs();
  f64x4 acc[2][2];
#pragma unroll
  for (int s = 0; s < 2; ++s) acc[s][0] = acc[s][1] = f64x4{0.0, 0.0, 0.0, 0.0};
  int bad = 0;
  const int chunks = (n + kARows - 1) / kARows;
  for (int c = 0; c < chunks; ++c) {
    const int rows = min(kARows, n - c * kARows);
    const double* xs = S.xs[buf];
    double pf[kSlots];
    stream_load(C, X, (c + 1 < chunks ? c + 1 : 0) * kARows, pf);
    if (wave < 4) {
      const f64x4 u = gate_scores(S, xs, d, wave);
      const double z = S.theta[lo][d];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * wave + hi + 4 * r;
        double q0 = 0.0, q1 = 0.0;
        if (row < rows) {
          const double score = (u[r] - z) * invT;
          const double pi = isfinite(score) ? fast_gate(score) : __builtin_nan("");
          const bool ok = isfinite(pi) && pi >= 0.0 && pi <= 1.0;
          bad |= !ok;
          if (ok) {
            q0 = fmax(1.0 - pi, floor);
            q1 = fmax(pi, floor);
          }
        }
        unsigned w3[3];
        split3_pair((float)q0, (float)q1, neg, w3[0], w3[1], w3[2]);
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
          S.u.fitb.QB[pl][2 * lo][row] = (unsigned short)(w3[pl] & 0xffffu);
          S.u.fitb.QB[pl][2 * lo + 1][row] = (unsigned short)(w3[pl] >> 16);
        }
      }
    } else {
      const int t = tid - 256, row = t >> 2;
      const double* xr = xs + row * kAStr;
      for (int f = t & 3; f < kFStr; f += 4) {
        double v = 0.0;
        if (f == p) {
          v = xr[d];
        } else if (f < p) {
          const int a = S.fa[f], b = S.fb[f];
          v = a >= 0 ? xr[a] : 1.0;
          if (b >= 0) v *= xr[b];
        }
        S.u.fitb.F[row][f] = v;
      }
    }
    __syncthreads();
    if (nw == 2) bf16_chunk<2>(S, lo, hi, oa, ob, neg, acc);
    else if (nw == 1) bf16_chunk<1>(S, lo, hi, oa, ob, neg, acc);
    stream_store(X, S.xs[buf ^ 1], pf);
    __syncthreads();
    buf ^= 1;
  }
  // The systems overwrite F and QB: every wave passed the loop's last barrier.
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int t = 16 * (wave + 8 * s) + lo;
    if (s < nw && t < ntri)
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) S.u.G[16 * mt + 4 * hi + r][t] = acc[s][mt][r];
  }
  if (wave < 4) {
    bad |= __shfl_xor(bad, 16, kWave);
    bad |= __shfl_xor(bad, 32, kWave);
    if (hi == 0) S.badw[wave][lo] = bad;
  }
  __syncthreads();
  if (tid < kAR) S.bad[tid] = S.badw[0][tid] | S.badw[1][tid] | S.badw[2][tid] | S.badw[3][tid];
  __syncthreads();
  return buf;
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
__device__ __attribute__((noinline)) void anneal_solve(AnnealCtx C, AnnealShared& S) {
  const int g = threadIdx.x >> 5, i = threadIdx.x & 31;
  const int p = C.p;
  for (int round = 0; round < 2 * kAR / kAGroups; ++round) {
    const int sys = round * kAGroups + g;
    const int r = sys >> 1, side = sys & 1;
    double* M = S.u.G[sys];
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
__device__ __attribute__((noinline)) int anneal_stats(AnnealCtx C, AnnealShared& S, int buf,
                                                      double T) {
  Stream X;
  stream_init(C, X);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lo = lane & 15, hi = lane >> 4;
  const int d = C.d, p = C.p, n = C.n;
  const double inv = 1.0 / (n * T), invT = 1.0 / T;
  double loss = 0.0, gz = 0.0, cnt = 0.0;  // start lo, this lane's rows
  f64x4 g = {0.0, 0.0, 0.0, 0.0};          // gradient tile: start hi + 4 r, dim lo
  const int chunks = (n + kARows - 1) / kARows;
  for (int c = 0; c < chunks; ++c) {
    const int rows = min(kARows, n - c * kARows);
    const double* xs = S.xs[buf];
    double pf[kSlots];
    stream_load(C, X, (c + 1 < chunks ? c + 1 : 0) * kARows, pf);
    if (wave < 4) {
      const f64x4 u = gate_scores(S, xs, d, wave);
      const double z = S.theta[lo][d];
#pragma unroll
      for (int r = 0; r < 4; ++r) S.u.st.pi[16 * wave + hi + 4 * r][lo] = fast_gate((u[r] - z) * invT);
    } else {
      // Predictions of rows 16 rt .. 16 rt + 15 for all 32 start-sides.
      const int rt = wave - 4;
      const double* xr = xs + (16 * rt + lo) * kAStr;
      f64x4 P0 = {0.0, 0.0, 0.0, 0.0}, P1 = {0.0, 0.0, 0.0, 0.0};
      for (int ks = 0; ks < (p + 3) / 4; ++ks) {
        const int f = 4 * ks + hi;
        double a = 0.0, b0 = 0.0, b1 = 0.0;
        if (f < p) {
          const int fa = S.fa[f], fb = S.fb[f];
          a = fa >= 0 ? xr[fa] : 1.0;
          if (fb >= 0) a *= xr[fb];
          b0 = S.beta[lo >> 1][lo & 1][f];
          b1 = S.beta[8 + (lo >> 1)][lo & 1][f];
        }
        P0 = mfma4(a, b0, P0);
        P1 = mfma4(a, b1, P1);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * rt + hi + 4 * r;
        S.u.st.P[row][lo] = P0[r];
        S.u.st.P[row][16 + lo] = P1[r];
      }
    }
    __syncthreads();
    // Wave w: rows 8 w .. 8 w + 7 as the gradient's 2 k-steps; the lane's
    // residual term is its own MFMA operand.
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int row = 8 * wave + 4 * j + hi;
      double cv = 0.0;
      if (row < rows) {
        const double pi = S.u.st.pi[row][lo];
        const double yv = xs[row * kAStr + d];
        double l = yv - S.u.st.P[row][2 * lo], rq = yv - S.u.st.P[row][2 * lo + 1];
        l *= l;
        rq *= rq;
        cv = (rq - l) * pi * (1.0 - pi) * inv;
        loss += (1.0 - pi) * l + pi * rq;
        gz -= cv;
        cnt += pi >= 0.5 ? 1.0 : 0.0;
      }
      const double xv = lo < d ? xs[row * kAStr + lo] : 0.0;
      g = mfma4(cv, xv, g);
    }
    stream_store(X, S.xs[buf ^ 1], pf);
    __syncthreads();
    buf ^= 1;
  }
  loss += __shfl_xor(loss, 16, kWave);
  loss += __shfl_xor(loss, 32, kWave);
  gz += __shfl_xor(gz, 16, kWave);
  gz += __shfl_xor(gz, 32, kWave);
  cnt += __shfl_xor(cnt, 16, kWave);
  cnt += __shfl_xor(cnt, 32, kWave);
#pragma unroll
  for (int r = 0; r < 4; ++r) S.u.red[wave][hi + 4 * r][lo] = g[r];
  if (hi == 0) {
    S.u.red[wave][lo][kADim] = loss;
    S.u.red[wave][lo][kADim + 1] = gz;
    S.u.red[wave][lo][kADim + 2] = cnt;
  }
  __syncthreads();
  if (tid < kAR * kADim) {
    const int r = tid >> 4, j = tid & 15;
    double v = 0.0;
    for (int w = 0; w < 8; ++w) v += S.u.red[w][r][j];
    S.gsum[r][j] = v;
  } else if (tid < kAR * kADim + 3 * kAR) {
    const int e = tid - kAR * kADim, r = e / 3, k = e - 3 * r;
    double v = 0.0;
    for (int w = 0; w < 8; ++w) v += S.u.red[w][r][kADim + k];
    S.acc[r][k] = v;
  }
  __syncthreads();
  if (tid < kAR) {
    const int r = tid;
    const double* w = S.theta[r];
    const double gzr = S.acc[r][1];
    double ww = 0.0, wg = 0.0;
    for (int j = 0; j < d; ++j) {
      ww += w[j] * w[j];
      wg += w[j] * S.gsum[r][j];
    }
    const double proj = wg / ww;
    double pp = 0.0, wp = 0.0;
    for (int j = 0; j < d; ++j) {
      const double v = -S.gsum[r][j] + proj * w[j];
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
  return buf;
}

// [SSE_L, SSE_R, n_right, bad] of every start's hard partition and fit; a
// wave per 2 starts, a lane per row (runs once, after the last iteration).
__device__ __attribute__((noinline)) void anneal_hard_stats(AnnealCtx C, AnnealShared& S) {
  Stream X;
  stream_init(C, X);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int d = C.d, p = C.p, n = C.n;
  double sl[2] = {0.0, 0.0}, sr[2] = {0.0, 0.0}, cnt[2] = {0.0, 0.0};
  int bad[2] = {0, 0};
  double* xs = S.xs[0];
  for (int base = 0; base < n; base += kARows) {
    const int rows = min(kARows, n - base);
    __syncthreads();
    stage_chunk(C, X, xs, base);
    __syncthreads();
    if (lane < rows) {
      const double* xr = xs + lane * kAStr;
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

__global__ __launch_bounds__(kAThreads, 4) void anneal_kernel(art_split_args args) {
  extern __shared__ __align__(16) unsigned char anneal_smem[];
  AnnealShared& S = *reinterpret_cast<AnnealShared*>(anneal_smem);
  const art_split_args& a = args;
  const int tid = threadIdx.x;
  const int q0 = blockIdx.x * kAR;
  const int node = a.node[q0];
  const int n = a.n, d = a.d, p = a.basis.p;
  AnnealCtx C;
  C.X = (gf64*)(a.X + (size_t)node * n * d);
  C.y = (gf64*)(a.y + (size_t)node * n);
  C.n = n;
  C.d = d;
  C.p = p;
  C.ridge = a.ridge;
  C.floor = a.floor;
  if (tid < kAR * (d + 1)) {
    const int r = tid / (d + 1), j = tid - r * (d + 1);
    S.theta[r][j] = a.theta0[(size_t)(q0 + r) * (d + 1) + j];
  }
  if (tid < p) {
    int fa, fb;
    basis_factors(a.basis, tid, fa, fb);
    S.fa[tid] = (signed char)fa;
    S.fb[tid] = (signed char)fb;
  }
  {
    Stream X;
    stream_init(C, X);
    stage_chunk(C, X, S.xs[0], 0);  // the first pass's first chunk
  }
  int buf = 0;
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
    if (a.child_precision == ART_PRECISION_MIXED) buf = anneal_fit_bf16(C, S, buf, T, C.floor);
    else buf = anneal_fit(C, S, buf, 0, T, C.floor);
    anneal_solve(C, S);
    buf = anneal_stats(C, S, buf, T);
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
    buf = anneal_fit(C, S, buf, 1, 1.0, 0.0);
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
  if (d < 1 || d > kADim || p < 1 || p > kAP || args->n < 1)
    return fail(ART_ERR_ARG, "art_split_anneal_batched: needs 1 <= d <= 16 and at most 17 "
                             "basis features");
  if (args->anneal_iters < 1 || !(args->t0 > 0.0) || !(args->t1 > 0.0))
    return fail(ART_ERR_ARG, "art_split_anneal_batched: needs iterations and positive temperatures");
  if (args->row_offsets)
    return fail(ART_ERR_ARG, "art_split_anneal_batched: nodes must have n rows each");
  if (args->child_precision != ART_PRECISION_FP64 && args->child_precision != ART_PRECISION_MIXED)
    return fail(ART_ERR_ARG, "art_split_anneal_batched: child_precision must be FP64 or MIXED");
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
