// Soft-oblique split optimizer: one persistent workgroup per problem
// (node x restart) runs the whole projected-gradient loop on the device.
//
// Hot-path pieces a4-a6: soft objective, gradient and tangent direction
// (art/objectives.py:132-340, art/splitters.py:71-98), the frozen-model Armijo
// search with AdaptiveAlpha and its retry policy (art/optimizers.py:20-121,
// art/splitters.py:400-460), commit/retraction, stop reasons, the
// last-feasible min-side fallback and the hard refit
// (art/splitters.py:461-690, art/_cuda_batched_optimizer.py:390-526).  The
// reference drives these from the host with a device->host transfer every
// iteration and every backtrack batch (art/_cuda_batched_optimizer.py:323,360);
// here the control flow, the AdaptiveAlpha state and every history live on
// the device, and one launch runs all problems of a cohort to completion.
//
// Per iteration a workgroup makes three passes over its node's rows:
//   fit       both sides' augmented Grams [phi | y]^T diag(q) [phi | y] by
//             v_mfma_f64_16x16x4 (lane (col, kq) builds feature 16t+col of row
//             4s+kq), then trace-scaled ridge, Cholesky and solve in LDS;
//   stats     one thread per row: residuals, loss, gradient coefficients and
//             X^T c (block reductions);
//   search    one thread per row evaluates H Armijo candidates at once with
//             the frozen residuals recomputed from the two coefficient vectors.
// Layout: X (K, n, d), y (K, n) float64 row-major; one node per problem index.

#include "art_common.h"
#include "art_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <new>

namespace art {

constexpr int kSplitMaxP = 32;
constexpr int kSplitMaxD = 31;
constexpr int kSplitMaxH = 8;  // candidates per trial pass
constexpr int kSplitScalars = 6;

// Stop reasons (strings in the Python layer; order fixed by the C ABI).
enum StopReason {
  kMaxIters = 0,
  kGradientTolerance = 1,
  kNonDescent = 2,
  kLineSearchFailed = 3,
  kMaxLineSearchFailures = 4,
  kAlphaMinReached = 5,
  kCholeskyFailure = 6,
  kMinSidePoints = 7,
  kNonfiniteSplit = 8,
};

__device__ inline double gate(double score) {
  // 1 / (1 + e) by v_rcp_f64 and two Newton steps (within an ulp of the
  // quotient; e + 1 lies in [1, 5.2e21], so the reciprocal is normal).
  const double v = 1.0 + exp(-fmin(50.0, fmax(-50.0, score)));
  double r = __builtin_amdgcn_rcp(v);
  r = fma(r, fma(-v, r, 1.0), r);
  return fma(r, fma(-v, r, 1.0), r);
}

// Workgroup and staged-chunk geometry.  Rows are staged in LDS with an odd
// stride of at least d + 1 doubles (x_0..x_{d-1}, y; odd, so a column read
// by consecutive rows spreads over the banks), as many rows per chunk as the
// table holds, up to one row per thread in the per-row phases.  One feature
// tile (at most 16 non-constant child features, d <= 16): 16 waves and up
// to 1024 rows (1024 at d <= 8, 512 at d = 16), so a node's pass is a few
// chunks with four waves per SIMD to hide each other's latency; two tiles
// (d <= 31): 4 waves and 128 rows.
template <int T> struct ChunkShape;
template <> struct ChunkShape<1> {
  static constexpr int waves = 8, maxrows = 512, table = 512 * 9, maxd = 16;
};
template <> struct ChunkShape<2> {
  static constexpr int waves = 4, maxrows = 128, table = 128 * 33, maxd = 31;
};
template <int T> constexpr int kWavesT = ChunkShape<T>::waves;
template <int T> constexpr int kThreadsT = 64 * ChunkShape<T>::waves;
template <int T> constexpr int kMaxRowsT = ChunkShape<T>::maxrows;

__host__ __device__ inline int row_stride(int d) { return (d + 1) | 1; }
template <int T>
__host__ __device__ inline int chunk_rows(int d) {
  const int r = ChunkShape<T>::table / row_stride(d) / 64 * 64;
  return r < ChunkShape<T>::maxrows ? r : ChunkShape<T>::maxrows;
}

template <int T>
struct SplitShared {
  double theta[kSplitMaxD + 1];
  double pdir[kSplitMaxD + 1];
  double last[kSplitMaxD + 1];
  double beta[2][kSplitMaxP];
  double lam[2];
  int status[2];
  double geometry[4];
  double packet[7];
  double red[kWavesT<T> * 256];  // MFMA tile reduction; reused by block reductions
  double M[2][(kSplitMaxP + 1) * (kSplitMaxP + 2) / 2];
  double piv[2];
  double bc[2];
  int ired[kWavesT<T>];
  double dred[kWavesT<T> * 64];
  double tot[64];
  double xs[ChunkShape<T>::table];  // staged rows: x_0..x_{d-1}, y, stride row_stride(d)
  double qs[2][kMaxRowsT<T>];  // per-row side weights (fit) or gradient terms (stats)
  signed char fa[kSplitMaxP];    // feature f = x[fa] * x[fb]; -1 stands for 1
  signed char fb[kSplitMaxP];
  signed char core[kSplitMaxP];  // basis column of non-constant feature k
  int nf;                        // number of non-constant features
  int cbias;                     // basis column of the constant, or -1
  int part_dead;                 // a multi-part node's exchange gave up waiting
};

struct SplitCtx {
  const double* X;
  const double* y;
  int n, d, p;
  int rows, sd;  // chunk rows and staged row stride
  int lo, hi;    // this workgroup's rows of the node: [lo, hi)
  Basis B;
  double T;
  double ridge;
};

// A node's rows split over `parts` workgroups (part 0..parts-1, consecutive
// block indices).  Every pass's row sums are exchanged through `buf` (two
// alternating slots of `stride` doubles per part and problem, stored and
// read at device scope so that workgroups on other XCDs see them) after a
// node barrier on `count[prob]`; each part then adds the parts' sums in part
// order, so every part holds bitwise the same node state and takes the same
// control path.  The barrier's wait is bounded: a part that waits longer
// than kPartSpins polls marks the launch failed (`failed`), stops waiting
// at later exchanges and finishes with meaningless results.
struct PartSync {
  int parts, part, prob, stride;
  int* count;
  double* buf;
  int* failed;
  int epoch, dead;
};
constexpr long kPartSpins = 1L << 22;
constexpr int kSplitMaxParts = 8;
static_assert(ChunkShape<1>::table >= 2 * ((kSplitMaxP + 1) * (kSplitMaxP + 2) / 2) + 1 &&
                  ChunkShape<2>::table >= 2 * ((kSplitMaxP + 1) * (kSplitMaxP + 2) / 2) + 1,
              "the row table stages a node's two exchanged systems");

// Feature tables of the basis (once per kernel, before any pass).
template <int T>
__device__ inline void init_basis_tables(const Basis& B, SplitShared<T>& S) {
  if (threadIdx.x == 0) {
    int nf = 0, cb = -1;
    for (int f = 0; f < B.p; ++f) {
      int a, b;
      basis_factors(B, f, a, b);
      S.fa[f] = (signed char)a;
      S.fb[f] = (signed char)b;
      if (a < 0) cb = f;
      else S.core[nf++] = (signed char)f;
    }
    S.nf = nf;
    S.cbias = cb;
  }
  __syncthreads();
}

// Row staging: a chunk's x values are one contiguous block of rows * d
// doubles (rows are contiguous in X), loaded element-wise by the workgroup
// (coalesced; ceil(rows d / threads) loads per thread, two for d = 2), and its
// y values one per thread.  A thread's loads for the next chunk are issued
// before the current chunk is processed (register double buffering: the
// barriers wait on LDS only, not on these loads), so the memory round trip
// hides behind a chunk of compute.  S.xs holds x_0..x_{d-1}, y per row with
// stride C.sd.
template <int T> constexpr int kStageT = (ChunkShape<T>::table + kThreadsT<T> - 1) / kThreadsT<T>;

template <int T>
struct Staged {
  double x[kStageT<T>];
  double y;
};

template <int T>
__device__ inline void load_rows(const SplitCtx& C, int base, int rows, Staged<T>& v) {
  const int tid = threadIdx.x, nx = rows * C.d;
  const double* src = C.X + (size_t)base * C.d;
#pragma unroll
  for (int i = 0; i < kStageT<T>; ++i) {
    const int e = tid + kThreadsT<T> * i;
    v.x[i] = e < nx ? src[e] : 0.0;
  }
  v.y = tid < rows ? C.y[base + tid] : 0.0;
}

template <int T>
__device__ inline void store_rows(const SplitCtx& C, SplitShared<T>& S, int rows,
                                  const Staged<T>& v) {
  const int tid = threadIdx.x, d = C.d, nx = rows * d;
#pragma unroll
  for (int i = 0; i < kStageT<T>; ++i) {
    const int e = tid + kThreadsT<T> * i;
    if (e < nx) {
      const int r = e / d;
      S.xs[r * C.sd + (e - r * d)] = v.x[i];
    }
  }
  if (tid < rows) S.xs[tid * C.sd + d] = v.y;
}

// Iterates this workgroup's rows of the node [C.lo, C.hi) in staged chunks:
// body(rows) runs with chunk rows [base, base + rows) in S.xs while the next
// chunk's loads are in flight.
template <int T, class Body>
__device__ inline void for_each_chunk(const SplitCtx& C, SplitShared<T>& S, Body body) {
  const int kChunk = C.rows;
  Staged<T> v;
  load_rows(C, C.lo, min(kChunk, C.hi - C.lo), v);
  for (int base = C.lo; base < C.hi; base += kChunk) {
    const int rows = min(kChunk, C.hi - base);
    __syncthreads();  // the previous chunk is consumed
    store_rows(C, S, rows, v);
    const int next = base + kChunk;
    if (next < C.hi) load_rows(C, next, min(kChunk, C.hi - next), v);
    __syncthreads();
    body(rows);
  }
}

template <int T>
__device__ inline double staged_feature(const SplitShared<T>& S, const double* xr, int f) {
  const int a = S.fa[f], b = S.fb[f];
  double v = a >= 0 ? xr[a] : 1.0;
  if (b >= 0) v *= xr[b];
  return v;
}

// Sums of K values over the block; result in S.tot[0..K).  All threads call.
template <int K, int T>
__device__ inline void block_sums(const double (&v)[K], SplitShared<T>& S, int count = K) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (k < count) {
      const double s = wave_sum(v[k]);
      if (lane == 0) S.dred[wave * 64 + k] = s;
    }
  __syncthreads();
  if ((int)threadIdx.x < count) {
    double t = S.dred[threadIdx.x];
#pragma unroll
    for (int w = 1; w < kWavesT<T>; ++w) t += S.dred[64 * w + threadIdx.x];
    S.tot[threadIdx.x] = t;
  }
  __syncthreads();
}

// Node-wide sums of an LDS vector v[0..len) over the problem's parts; every
// part ends with the same sums.  All threads of every part call this at the
// same point of the control flow.
template <int T>
__device__ inline void parts_sum(PartSync& P, SplitShared<T>& S, double* v, int len) {
  if (P.parts == 1) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const size_t slot = ((size_t)P.prob * 2 + (P.epoch & 1)) * P.parts;
  double* mine = P.buf + (slot + P.part) * P.stride;
  __syncthreads();  // v is complete
  for (int k = tid; k < len; k += nt) store_dev(&mine[k], v[k]);
  // Every thread's stores complete before the barrier, so before the count.
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) {
    int* cnt = P.count + P.prob;
    __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int target = P.parts * (P.epoch + 1);
    long spins = 0;
    while (!P.dead && __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(4);
      if (++spins > kPartSpins) {
        P.dead = 1;
        __hip_atomic_store(P.failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    S.part_dead = P.dead;
  }
  __syncthreads();
  P.dead = S.part_dead;
  ++P.epoch;
  const double* all = P.buf + slot * P.stride;
  for (int k = tid; k < len; k += nt) {
    double s = 0.0;
    for (int w = 0; w < P.parts; ++w) s += load_dev(&all[(size_t)w * P.stride + k]);
    v[k] = s;
  }
  __syncthreads();
}

// Soft (hard = false) or hard (hard = true) weighted fits of both sides at
// the current theta; half = true fits with pi = 1/2 (the parent fit).
// Writes beta, lam and status (0 / -1 bad gate / pivot / -2 non-finite /
// -3 solve) for both sides into shared state.
//
// Per staged chunk of rows: one thread per row computes the gate and both
// side weights q_L = max(1 - pi, floor), q_R = max(pi, floor) and the
// weighted sums of 1, y and y^2; then v_mfma_f64_16x16x4 accumulates the
// Gram of the non-constant features (lane (col, kq) feeds feature 16t + col
// of row 4s + kq), and the same lanes accumulate the feature sums against
// q and q y (the constant and y columns of the augmented system).
// Child systems up to kWaveSolveP wide are solved by one wave each.
constexpr int kWaveSolveP = 16;

// Trace-scaled ridge, Cholesky and solve of one packed child system M
// ((p+1) x (p+1): G and the row h) by one wave, rows in registers; the same
// status contract as the workgroup form below (-1 bad gates, -2 non-finite,
// k+1 first pivot <= 0, -3 non-finite solution).  col: 64 doubles of LDS for
// this wave.  Writes beta[0..p) and lam; returns the status (uniform).
template <int PMAX>
__device__ inline int wave_side_solve(double* M, int p, double ridge, int bad, int lane,
                                      double* col, double* beta, double& lam_out) {
  const double tr = wave_sum(lane < p ? M[tri(lane, lane)] : 0.0);
  const double lam = ridge == 0.0 ? 0.0 : ridge * (tr / p);
  int st = bad ? -1 : ((!isfinite(tr) || !isfinite(lam)) ? -2 : 0);
  double x = 0.0;
  if (st == 0) {
    // Each loop tests the lane against every column; a fresh opaque copy of
    // the lane index per loop keeps those tests out of scalar registers.
    double a[PMAX];
    {
      const int me = opaque_lane(lane);
#pragma unroll
      for (int j = 0; j < PMAX; ++j) a[j] = (j <= me && me < p) ? M[tri(me, j)] : 0.0;
    }
    const double h = lane < p ? M[tri(p, lane)] : 0.0;
#pragma unroll
    for (int j = 0; j < PMAX; ++j)
      if (j == lane) a[j] += lam;
    double diag;
    st = wave_cholesky<PMAX>(a, p, lane, diag, col);
    if (st == 0) {
      const int me = opaque_lane(lane);
      int nf = me < p && !isfinite(h);
#pragma unroll
      for (int j = 0; j < PMAX; ++j) nf |= (j <= me && me < p && !isfinite(a[j]));
      if (wave_or(nf)) st = -2;
    }
    if (st == 0) {
      const double dinv = 1.0 / diag;
      const double z = wave_forward<PMAX>(a, dinv, p, lane, h);
      // The factor's rows go back to M, from which every lane reads its
      // column for the back substitution (one wave: its LDS operations
      // complete in order; the fence keeps the reads after the writes).
      const int me = opaque_lane(lane);
#pragma unroll
      for (int j = 0; j < PMAX; ++j)
        if (j <= me && me < p) M[tri(me, j)] = a[j];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const int me2 = opaque_lane(lane);
#pragma unroll
      for (int k = 0; k < PMAX; ++k) a[k] = (k > me2 && k < p) ? M[tri(k, me2)] : 0.0;
      x = wave_backward<PMAX>(a, dinv, p, lane, z);
      if (wave_or(lane < p && !isfinite(x))) st = -3;
    }
  }
  if (lane < p) beta[lane] = st == 0 ? x : __builtin_nan("");
  if (lane == 0) lam_out = lam;
  return st;
}

template <int T>
__device__ void fit_sides(const SplitCtx& C, SplitShared<T>& S, PartSync& P, bool hard,
                          double floor, bool half = false) {
  constexpr int NP = T * (T + 1) / 2;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, kq = lane >> 4;
  const int d = C.d, p = C.p, n = C.n, sd = C.sd, nf = S.nf;
  int fa[T], fb[T];
  bool live[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int k = 16 * t + col;
    live[t] = k < nf;
    const int f = live[t] ? S.core[k] : 0;
    fa[t] = live[t] ? S.fa[f] : 0;
    fb[t] = live[t] ? S.fb[f] : -1;
  }
  f64x4 acc0[NP], acc1[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    acc0[q] = f64x4{0.0, 0.0, 0.0, 0.0};
    acc1[q] = f64x4{0.0, 0.0, 0.0, 0.0};
  }
  double e[4 * T];  // per tile: sum q0 v, q0 y v, q1 v, q1 y v
#pragma unroll
  for (int k = 0; k < 4 * T; ++k) e[k] = 0.0;
  double sc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // sum q0, q0 y, q0 y^2, q1, ...
  int bad = 0;
  const double z = S.theta[d];
  for_each_chunk(C, S, [&](int rows) {
    if (tid < C.rows) {
      double q0 = 0.0, q1 = 0.0;
      if (tid < rows) {
        const double* xr = S.xs + tid * sd;
        double u = 0.0;
        for (int j = 0; j < d; ++j) u = fma(xr[j], S.theta[j], u);
        double pi;
        if (half) {
          pi = 0.5;
        } else if (hard) {
          pi = (isfinite(u) && isfinite(z)) ? (double)(u >= z) : __builtin_nan("");
        } else {
          const double score = (u - z) / C.T;
          pi = (C.T > 0.0 && isfinite(C.T) && isfinite(score)) ? gate(score) : __builtin_nan("");
        }
        const bool okg = isfinite(pi) && pi >= 0.0 && pi <= 1.0;
        bad |= !okg;
        q0 = okg ? fmax(1.0 - pi, floor) : 0.0;
        q1 = okg ? fmax(pi, floor) : 0.0;
        const double yv = xr[d];
        sc[0] += q0;
        sc[1] += q0 * yv;
        sc[2] += q0 * yv * yv;
        sc[3] += q1;
        sc[4] += q1 * yv;
        sc[5] += q1 * yv * yv;
      }
      S.qs[0][tid] = q0;
      S.qs[1][tid] = q1;
    }
    __syncthreads();
    for (int st = wave; st < C.rows / 4; st += kWavesT<T>) {
      const int r = 4 * st + kq;
      const bool valid = r < rows;
      const double* xr = S.xs + r * sd;
      const double q0 = S.qs[0][r], q1 = S.qs[1][r];
      const double yv = valid ? xr[d] : 0.0;
      double v[T];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        double x = 0.0;
        if (valid && live[t]) {
          x = xr[fa[t]];
          if (fb[t] >= 0) x *= xr[fb[t]];
        }
        v[t] = x;
      }
      int q = 0;
#pragma unroll
      for (int t1 = 0; t1 < T; ++t1)
#pragma unroll
        for (int t2 = t1; t2 < T; ++t2, ++q) {
          acc0[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(q0 * v[t1], v[t2], acc0[q], 0, 0, 0);
          acc1[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(q1 * v[t1], v[t2], acc1[q], 0, 0, 0);
        }
#pragma unroll
      for (int t = 0; t < T; ++t) {
        e[4 * t + 0] += q0 * v[t];
        e[4 * t + 1] += q0 * yv * v[t];
        e[4 * t + 2] += q1 * v[t];
        e[4 * t + 3] += q1 * yv * v[t];
      }
    }
  });
  // Assemble the packed augmented systems [phi | y]^T diag(q) [phi | y].
  const int nel = (p + 1) * (p + 2) / 2;
  __syncthreads();
  for (int k = tid; k < nel; k += blockDim.x) {
    S.M[0][k] = 0.0;
    S.M[1][k] = 0.0;
  }
  // Gram of the non-constant features: element (a, b) of tile pair (t1, t2)
  // is G[16 t1 + a][16 t2 + b]; keep a <= b on diagonal tiles.
  for (int side = 0; side < 2; ++side) {
    int q = 0;
#pragma unroll
    for (int t1 = 0; t1 < T; ++t1)
#pragma unroll
      for (int t2 = t1; t2 < T; ++t2, ++q) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r)
          S.red[wave * 256 + (kq + 4 * r) * 16 + col] = side == 0 ? acc0[q][r] : acc1[q][r];
        __syncthreads();
        double v = 0.0;
        if (tid < 256) {
          v = S.red[tid];
#pragma unroll
          for (int w = 1; w < kWavesT<T>; ++w) v += S.red[256 * w + tid];
        }
        const int ka = 16 * t1 + (tid >> 4), kb = 16 * t2 + (tid & 15);
        if (tid < 256 && ka < nf && kb < nf && (t1 != t2 || ka <= kb)) {
          const int ca = S.core[ka], cb = S.core[kb];
          S.M[side][ca >= cb ? tri(ca, cb) : tri(cb, ca)] = v;
        }
      }
  }
  // Feature sums against q and q y: reduce over the four row lanes (kq) of
  // each wave, then over the waves.
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4 * T; ++k) {
    double s = e[k];
    s += __shfl_xor(s, 16, kWave);
    s += __shfl_xor(s, 32, kWave);
    if (kq == 0) S.red[wave * 256 + k * 16 + col] = s;
  }
  __syncthreads();
  if (tid < 64 * T) {
    const int k = tid >> 4, c = tid & 15;
    const int t = k >> 2, kind = k & 3;
    const int feat = 16 * t + c;
    double s = S.red[tid];
#pragma unroll
    for (int w = 1; w < kWavesT<T>; ++w) s += S.red[256 * w + tid];
    if (feat < nf) {
      const int cf = S.core[feat];
      const int side = kind >> 1;
      if (kind & 1) {
        S.M[side][tri(p, cf)] = s;
      } else if (S.cbias >= 0) {
        const int cb = S.cbias;
        S.M[side][cf >= cb ? tri(cf, cb) : tri(cb, cf)] = s;
      }
    }
  }
  block_sums<6>(sc, S);
  if (tid == 0) {
    for (int side = 0; side < 2; ++side) {
      if (S.cbias >= 0) {
        S.M[side][tri(S.cbias, S.cbias)] = S.tot[3 * side];
        S.M[side][tri(p, S.cbias)] = S.tot[3 * side + 1];
      }
      S.M[side][tri(p, p)] = S.tot[3 * side + 2];
    }
  }
  __syncthreads();
  bad = block_or(bad, S.ired);
  if (P.parts > 1) {
    // The node's systems from every part's rows, staged contiguously in the
    // row table (free between passes; at least 2 nel + 1 doubles).
    double* v = S.xs;
    for (int k = tid; k < nel; k += blockDim.x) {
      v[k] = S.M[0][k];
      v[nel + k] = S.M[1][k];
    }
    if (tid == 0) v[2 * nel] = bad;
    parts_sum(P, S, v, 2 * nel + 1);
    for (int k = tid; k < nel; k += blockDim.x) {
      S.M[0][k] = v[k];
      S.M[1][k] = v[nel + k];
    }
    bad = v[2 * nel] != 0.0;
    __syncthreads();
  }
  if (p <= kWaveSolveP) {
    // Small child systems: wave `side` factors and solves side `side` in
    // registers (no barrier per column), both sides at once.
    // Width 8 for p <= 8: the same results (steps past p change nothing),
    // half the column and substitution steps.
    if (wave < 2) {
      const int side = wave;
      S.status[side] =
          p <= 8 ? wave_side_solve<8>(S.M[side], p, C.ridge, bad, lane, S.red + 64 * side,
                                      S.beta[side], S.lam[side])
                 : wave_side_solve<kWaveSolveP>(S.M[side], p, C.ridge, bad, lane,
                                                S.red + 64 * side, S.beta[side], S.lam[side]);
    }
    __syncthreads();
    return;
  }
  for (int side = 0; side < 2; ++side) {
    double* M = S.M[side];
    const double tr = block_sum(tid < p ? M[tri(tid, tid)] : 0.0, S.red);
    const double ridge = C.ridge;
    const double lam = ridge == 0.0 ? 0.0 : ridge * (tr / p);
    int st = bad ? -1 : ((!isfinite(tr) || !isfinite(lam)) ? -2 : 0);
    double x = 0.0;
    if (st == 0) {
      if (tid < p) M[tri(tid, tid)] += lam;
      __syncthreads();
      st = rows_cholesky(M, &S.piv[side], p, tid);
      if (st == 0) {
        int nfin = 0;
        if (tid < p) {
          for (int j = 0; j <= tid; ++j) nfin |= !isfinite(M[tri(tid, j)]);
          nfin |= !isfinite(M[tri(p, tid)]);
        }
        if (block_or(nfin, S.ired)) st = -2;
      }
      if (st == 0) {
        x = rows_solve(M, S.bc, p, tid, tid < p ? M[tri(p, tid)] : 0.0);
        if (block_or(tid < p && !isfinite(x), S.ired)) st = -3;
      }
    }
    if (tid < p) S.beta[side][tid] = st == 0 ? x : __builtin_nan("");
    if (tid == 0) {
      S.status[side] = st;
      S.lam[side] = lam;
    }
    __syncthreads();
  }
}

// Predictions of both sides for one staged row.
template <int T>
__device__ inline void side_predictions(const SplitShared<T>& S, const double* xr, int p, double& pl,
                                        double& pr) {
  pl = 0.0;
  pr = 0.0;
  for (int f = 0; f < p; ++f) {
    const double v = staged_feature(S, xr, f);
    pl = fma(v, S.beta[0][f], pl);
    pr = fma(v, S.beta[1][f], pr);
  }
}

// Residual statistics at the current fit; then the tangent direction,
// geometry and control packet (art/_cuda_optimizer_kernels.cu:108-195):
// packet = [loss, |projected grad|, slope, n_right, lambda_L, lambda_R, bad].
// The gradient X^T c is formed per chunk from the per-row terms c in LDS,
// a (column, row group) thread grid over the staged rows.
template <int T>
__device__ void stats_and_direction(const SplitCtx& C, SplitShared<T>& S, PartSync& P) {
  const int tid = threadIdx.x;
  const int d = C.d, p = C.p, n = C.n, sd = C.sd;
  const double z = S.theta[d];
  const double inv = 1.0 / (n * C.T);
  const int gj = tid & 31, gg = tid >> 5, groups = (int)(blockDim.x >> 5);
  double sums[3] = {0.0, 0.0, 0.0};  // loss, -sum c, count(pi >= 1/2)
  double g = 0.0;                   // partial of (X^T c)[gj]
  for_each_chunk(C, S, [&](int rows) {
    if (tid < C.rows) {
      double c = 0.0;
      if (tid < rows) {
        const double* xr = S.xs + tid * sd;
        double pl, pr;
        side_predictions(S, xr, p, pl, pr);
        double u = 0.0;
        for (int j = 0; j < d; ++j) u = fma(xr[j], S.theta[j], u);
        const double pi = gate((u - z) / C.T);
        const double yv = xr[d];
        double l = yv - pl, rr = yv - pr;
        l *= l;
        rr *= rr;
        c = (rr - l) * pi * (1.0 - pi) * inv;
        sums[0] += (1.0 - pi) * l + pi * rr;
        sums[1] -= c;
        sums[2] += pi >= 0.5 ? 1.0 : 0.0;
      }
      S.qs[0][tid] = c;
    }
    __syncthreads();
    if (gj < d)
      for (int r = gg; r < rows; r += groups) g = fma(S.qs[0][r], S.xs[r * sd + gj], g);
  });
  block_sums<3>(sums, S);
  __syncthreads();
  S.red[tid] = g;
  __syncthreads();
  if (tid < d) {
    double s = 0.0;
    for (int q = 0; q < groups; ++q) s += S.red[q * 32 + tid];
    S.tot[3 + tid] = s;
  }
  __syncthreads();
  parts_sum(P, S, S.tot, 3 + d);
  const double loss = S.tot[0] / n, gz = S.tot[1], cnt = S.tot[2];
  if (tid == 0) {
    const double* w = S.theta;
    const double* gw = S.tot + 3;
    double ww = 0.0, wg = 0.0;
    for (int j = 0; j < d; ++j) {
      ww += w[j] * w[j];
      wg += w[j] * gw[j];
    }
    const double proj = wg / ww;
    double pp = 0.0, wp = 0.0, gp = 0.0;
    for (int j = 0; j < d; ++j) {
      const double v = -gw[j] + proj * w[j];
      S.pdir[j] = v;
      pp += v * v;
      wp += w[j] * v;
      gp += gw[j] * v;
    }
    const double norm = sqrt(pp + gz * gz);
    S.pdir[d] = -gz;
    S.geometry[0] = ww;
    S.geometry[1] = wp;
    S.geometry[2] = pp;
    S.geometry[3] = -gz;
    S.packet[0] = loss;
    S.packet[1] = norm;
    S.packet[2] = gp - gz * gz;
    S.packet[3] = cnt;
    S.packet[4] = S.lam[0];
    S.packet[5] = S.lam[1];
    const bool failed = S.status[0] != 0 || S.status[1] != 0;
    S.packet[6] = (failed || !(ww > 0.0) || !isfinite(ww) || !isfinite(norm) ||
                   !isfinite(S.packet[2]) || !isfinite(loss))
                      ? 1.0
                      : 0.0;
  }
  __syncthreads();
}

// Frozen-model trial losses of `count` (<= kSplitMaxH) candidates alphas[0..count).
template <int T>
__device__ void trial_losses(const SplitCtx& C, SplitShared<T>& S, PartSync& P,
                             const double* alphas, int count, double* out) {
  const int tid = threadIdx.x;
  const int d = C.d, p = C.p, n = C.n, sd = C.sd;
  const double z = S.theta[d], temp = C.T;
  const double g0 = S.geometry[0], g1 = S.geometry[1], g2 = S.geometry[2], g3 = S.geometry[3];
  double f[kSplitMaxH];
  double rinv[kSplitMaxH];
  bool okc[kSplitMaxH];
#pragma unroll
  for (int k = 0; k < kSplitMaxH; ++k) {
    f[k] = 0.0;
    const double al = k < count ? alphas[k] : 0.0;
    const double r2 = g0 + 2.0 * al * g1 + al * al * g2;
    okc[k] = (r2 > 0.0) && isfinite(r2) && (temp > 0.0) && isfinite(temp);
    rinv[k] = okc[k] ? 1.0 / sqrt(r2) : 0.0;
  }
  for_each_chunk(C, S, [&](int rows) {
    if (tid < rows) {
      const double* xr = S.xs + tid * sd;
      double pl, pr;
      side_predictions(S, xr, p, pl, pr);
      const double yv = xr[d];
      double l = yv - pl, rr = yv - pr;
      l *= l;
      rr *= rr;
      double u = 0.0, v = 0.0;
      for (int j = 0; j < d; ++j) {
        u = fma(xr[j], S.theta[j], u);
        v = fma(xr[j], S.pdir[j], v);
      }
#pragma unroll
      for (int k = 0; k < kSplitMaxH; ++k) {
        if (k < count) {
          const double al = alphas[k];
          const double score = ((u + al * v) * rinv[k] - z - al * g3) / temp;
          const double pi = gate(score);
          const double fk = (1.0 - pi) * l + pi * rr;
          f[k] += (okc[k] && isfinite(score) && isfinite(fk)) ? fk : INFINITY;
        }
      }
    }
  });
  block_sums<kSplitMaxH>(f, S, count);
  parts_sum(P, S, S.tot, count);
  if (tid < count) out[tid] = S.tot[tid] / n;
  __syncthreads();
}

// Hard partition statistics after a hard fit (art/_cuda_batched_optimizer.cu:92-120):
// [MSE, MSE_L, MSE_R, n_right, lambda_L, lambda_R, bad].
template <int T>
__device__ void hard_stats(const SplitCtx& C, SplitShared<T>& S, PartSync& P, bool refit,
                           double* out, bool all_right = false) {
  const int tid = threadIdx.x;
  const int d = C.d, p = C.p, n = C.n, sd = C.sd;
  const double z = S.theta[d];
  double sums[3] = {0.0, 0.0, 0.0};  // SSE left, SSE right, count right
  int bad = 0;
  for_each_chunk(C, S, [&](int rows) {
    if (tid < rows) {
      const double* xr = S.xs + tid * sd;
      double u = 0.0;
      for (int j = 0; j < d; ++j) u = fma(xr[j], S.theta[j], u);
      const double pi = all_right ? 1.0
                        : ((isfinite(u) && isfinite(z)) ? (double)(u >= z) : __builtin_nan(""));
      bad |= !isfinite(pi);
      const bool right = pi >= 0.5;
      if (refit) {
        const double* beta = S.beta[right ? 1 : 0];
        double pred = 0.0;
        for (int f = 0; f < p; ++f) pred = fma(staged_feature(S, xr, f), beta[f], pred);
        const double e = xr[d] - pred;
        sums[right ? 1 : 0] += e * e;
      }
      sums[2] += right ? 1.0 : 0.0;
    }
  });
  block_sums<3>(sums, S);
  bad = block_or(bad, S.ired);
  if (P.parts > 1) {
    if (tid == 0) S.tot[3] = bad;
    parts_sum(P, S, S.tot, 4);
    bad = S.tot[3] != 0.0;
  }
  if (tid == 0) {
    const double sl = S.tot[0], sr = S.tot[1], cnt = S.tot[2];
    out[3] = cnt;
    if (refit) {
      out[0] = (sl + sr) / n;
      out[1] = cnt < n ? sl / (n - cnt) : 0.0;
      out[2] = cnt > 0 ? sr / cnt : 0.0;
      out[4] = S.lam[0];
      out[5] = S.lam[1];
      out[6] = (S.status[0] != 0 || S.status[1] != 0 || !isfinite(out[0])) ? 1.0 : 0.0;
    } else {
      out[0] = out[1] = out[2] = 0.0;
      out[4] = out[5] = 0.0;
      out[6] = bad ? 1.0 : 0.0;
    }
  }
  __syncthreads();
}

__device__ inline double clip_alpha(const art_split_args* a, double v) {
  return fmin(fmax(v, a->alpha_min), a->alpha_max);
}

template <int T>
__global__ __launch_bounds__(kThreadsT<T>) void split_optimize_kernel(art_split_args args,
                                                                      PartSync P) {
  __shared__ SplitShared<T> S;
  __shared__ double alphas[kSplitMaxH];
  __shared__ double losses[kSplitMaxH];
  __shared__ int sflag[8];
  const int prob = blockIdx.x / P.parts;
  P.part = blockIdx.x - prob * P.parts;
  P.prob = prob;
  P.epoch = 0;
  P.dead = 0;
  const int tid = threadIdx.x;
  const art_split_args* a = &args;
  const int node = a->node[prob];
  const int d = a->d;
  const int64_t r0 = a->row_offsets ? a->row_offsets[node] : (int64_t)node * a->n;
  const int n = a->row_offsets ? (int)(a->row_offsets[node + 1] - r0) : a->n;
  SplitCtx C;
  C.X = a->X + (size_t)r0 * d;
  C.y = a->y + r0;
  C.n = n;
  C.d = d;
  C.B = a->basis;
  C.p = a->basis.p;
  C.T = a->temperature[node];
  C.ridge = a->ridge;
  C.sd = row_stride(d);
  C.rows = chunk_rows<T>(d);
  {  // this part's rows: whole chunks, as even a share as they allow
    const int chunks = (n + C.rows - 1) / C.rows;
    const int per = (chunks + P.parts - 1) / P.parts;
    C.lo = min(n, P.part * per * C.rows);
    C.hi = min(n, C.lo + per * C.rows);
  }
  const int p = C.p;
  const int MI = a->max_iters;
  double* hist = a->history + (size_t)prob * 5 * (MI + 1);
  if (tid <= d) S.theta[tid] = a->theta0[(size_t)prob * (d + 1) + tid];
  init_basis_tables(C.B, S);
  // Unit-normalize the start (art/splitters.py:71-83 is applied by the caller too).
  if (tid == 0) {
    double nw = 0.0;
    for (int j = 0; j < d; ++j) nw += S.theta[j] * S.theta[j];
    nw = sqrt(nw);
    for (int j = 0; j < d; ++j) S.theta[j] /= nw;
  }
  __syncthreads();

  if (a->mode == 2) {
    // Parent fit (art/_cuda_batched_optimizer.py:403-411): pi = 1/2 on both
    // sides is the unweighted fit (the factor cancels in G, h and the
    // trace-scaled ridge); its MSE over all rows is the parent loss.
    fit_sides<T>(C, S, P, false, 0.0, true);
    double* hard = a->hard + (size_t)prob * 8;
    hard_stats(C, S, P, true, hard, true);
    if (tid < p) {
      a->beta_out[((size_t)prob * 2 + 0) * p + tid] = S.beta[1][tid];
      a->beta_out[((size_t)prob * 2 + 1) * p + tid] = S.beta[1][tid];
    }
    if (tid == 0) {
      hard[7] = 0.0;
      int* info = a->info + (size_t)prob * 8;
      for (int k = 0; k < 8; ++k) info[k] = 0;
      info[0] = hard[6] != 0.0 ? kCholeskyFailure : kMaxIters;
      info[6] = hard[6] != 0.0;
    }
    return;
  }

  int reason = kMaxIters, converged = 0, it = 0, feasible_it = -1, error = 0;
  long physical = 0, logical = 0;
  double alpha_ctrl = clip_alpha(a, a->alpha0);
  const int required =
      a->row_offsets ? max(a->min_side_points, (int)ceil(a->min_side_fraction * n)) : a->required;
  const int S_ = a->max_backtracks + 1;
  const int H = a->H;

  if (a->mode == 1) {
    // Annealed fixed-step projected gradient: T_k geometric from t0 to t1.
    for (int k = 0; k < a->anneal_iters; ++k) {
      C.T = a->t0 * pow(a->t1 / a->t0, (double)k / (double)max(a->anneal_iters - 1, 1));
      fit_sides<T>(C, S, P, false, a->floor);
      stats_and_direction(C, S, P);
      if (tid == 0 && k <= MI) hist[k] = S.packet[0];
      if (S.packet[6] != 0.0) {
        error = kCholeskyFailure;
        break;
      }
      if (tid == 0) {
        const double st = a->step;
        const double r = sqrt(S.geometry[0] + 2.0 * st * S.geometry[1] + st * st * S.geometry[2]);
        for (int j = 0; j < d; ++j) S.theta[j] = (S.theta[j] + st * S.pdir[j]) / r;
        S.theta[d] += st * S.pdir[d];
      }
      __syncthreads();
      ++it;
    }
    C.T = a->temperature[node];
  } else {
    fit_sides<T>(C, S, P, false, a->floor);
    stats_and_direction(C, S, P);
    const double g0 = S.packet[1];
    const double tol = a->grad_atol + a->grad_rtol * g0;
    if (tid == 0) {
      hist[0 * (MI + 1) + 0] = S.packet[0];
      hist[1 * (MI + 1) + 0] = S.packet[1];
      a->scalars[(size_t)prob * kSplitScalars + 0] = g0;
      a->scalars[(size_t)prob * kSplitScalars + 1] = tol;
    }
    auto feasible_now = [&]() {
      return S.packet[6] == 0.0 && S.packet[3] >= required && n - S.packet[3] >= required;
    };
    if (feasible_now()) {
      if (tid <= d) S.last[tid] = S.theta[tid];
      feasible_it = 0;
    }
    __syncthreads();
    while (true) {
      if (S.packet[6] != 0.0) {
        error = kCholeskyFailure;
        break;
      }
      if (it >= MI) break;
      if (S.packet[1] <= tol) {
        converged = 1;
        reason = kGradientTolerance;
        break;
      }
      if (S.packet[2] >= 0.0) {
        reason = kNonDescent;
        break;
      }
      // Line search with retries (art/splitters.py:423-460).
      int failures = 0, accepted = 0, selected = -1;
      double step = 0.0, value = INFINITY;
      while (true) {
        // Candidates alpha_k = a0 rho^k by repeated multiplication
        // (art/_cuda_batched_optimizer.cu:47-57); a logical batch of H
        // candidates is evaluated in passes of at most kSplitMaxH.  With one
        // workgroup per problem (launches of many problems, bound by row
        // work) the first candidate, usually accepted, goes alone and the
        // rest follow only when it fails; with several parts per problem
        // (bound by the passes' latency) the batch goes at once.  The batch's
        // choice -- its first candidate that satisfies Armijo -- and every
        // loss are the same either way; `physical` counts the batch, as the
        // reference's batched search does.
        const int lead = P.parts == 1 ? 1 : kSplitMaxH;
        double acur = a->adaptive ? alpha_ctrl : a->alpha0;
        selected = -1;
        for (int offset = 0; offset < S_; offset += H) {
          const int count = min(H, S_ - offset);
          int sel = -1;
          for (int sub = 0; sub < count && sel < 0;) {
            const int c2 = min(sub == 0 ? lead : kSplitMaxH, count - sub);
            __syncthreads();  // every thread has finished reading the last pass
            if (tid == 0) {
              double t = acur;
              for (int k = 0; k < c2; ++k) {
                alphas[k] = t;
                t *= a->rho;
              }
            }
            for (int k = 0; k < c2; ++k) acur *= a->rho;
            __syncthreads();
            trial_losses(C, S, P, alphas, c2, losses);
            for (int k = 0; k < c2 && sel < 0; ++k) {
              value = losses[k];
              step = alphas[k];
              if (isfinite(value) && value <= S.packet[0] + a->armijo_c * step * S.packet[2])
                sel = offset + sub + k;
            }
            sub += c2;
          }
          physical += count;
          logical += sel >= 0 ? sel - offset + 1 : count;
          if (sel >= 0) {
            selected = sel;
            break;
          }
        }
        const bool ok = selected >= 0;
        if (a->adaptive) {
          if (!ok) alpha_ctrl = clip_alpha(a, step * a->rho);
          else if (selected == 0) alpha_ctrl = clip_alpha(a, step * a->grow);
          else if (selected >= a->heavy) alpha_ctrl = clip_alpha(a, step * a->recovery);
          else alpha_ctrl = clip_alpha(a, step / a->rho);
        }
        if (ok) {
          accepted = 1;
          break;
        }
        ++failures;
        if (!a->adaptive) {
          reason = kLineSearchFailed;
          break;
        }
        if (failures >= a->max_failures) {
          reason = kMaxLineSearchFailures;
          break;
        }
        if (alpha_ctrl <= a->alpha_min) {
          reason = kAlphaMinReached;
          break;
        }
      }
      if (!accepted) break;
      ++it;
      // Commit: w <- (w + alpha p_w) / r(alpha), z <- z + alpha p_z.
      __syncthreads();
      if (tid == 0) {
        const double r = sqrt(S.geometry[0] + 2.0 * step * S.geometry[1] +
                              step * step * S.geometry[2]);
        for (int j = 0; j < d; ++j) S.theta[j] = (S.theta[j] + step * S.pdir[j]) / r;
        S.theta[d] += step * S.pdir[d];
      }
      __syncthreads();
      fit_sides<T>(C, S, P, false, a->floor);
      stats_and_direction(C, S, P);
      if (tid == 0) {
        hist[0 * (MI + 1) + it] = S.packet[0];
        hist[1 * (MI + 1) + it] = S.packet[1];
        hist[2 * (MI + 1) + it - 1] = step;
        hist[3 * (MI + 1) + it - 1] = (double)selected;
        hist[4 * (MI + 1) + it - 1] = value;
      }
      if (feasible_now()) {
        if (tid <= d) S.last[tid] = S.theta[tid];
        feasible_it = it;
      }
      __syncthreads();
    }
  }

  // Final partition (art/_cuda_batched_optimizer.py:413-430).
  double* hard = a->hard + (size_t)prob * 8;
  int fallback = 0;
  const int finished = !error;  // the final partition stage runs
  if (!error) {
    const double right = S.packet[3];
    fallback = a->mode == 0 && a->fallback && feasible_it >= 0 &&
               fmin(right, n - right) < required;
    if (fallback && tid <= d) S.theta[tid] = S.last[tid];
    __syncthreads();
    if (tid == 0) {
      double nw = 0.0;
      for (int j = 0; j < d; ++j) nw += S.theta[j] * S.theta[j];
      nw = sqrt(nw);
      for (int j = 0; j < d; ++j) S.theta[j] /= nw;
    }
    __syncthreads();
    if (a->hard_refit) fit_sides<T>(C, S, P, true, 0.0);
    hard_stats(C, S, P, a->hard_refit != 0, hard);
    if (tid == 0) {
      const double right2 = hard[3];
      if (!a->hard_refit && hard[6] != 0.0) error = kNonfiniteSplit;
      else if (fmin(right2, n - right2) < required) error = kMinSidePoints;
      else if (hard[6] != 0.0) error = kCholeskyFailure;
      sflag[0] = error;
    }
    __syncthreads();
    error = sflag[0];
  }
  if (tid <= d) a->theta_out[(size_t)prob * (d + 1) + tid] = S.theta[tid];
  if (tid < p) {
    a->beta_out[((size_t)prob * 2 + 0) * p + tid] = S.beta[0][tid];
    a->beta_out[((size_t)prob * 2 + 1) * p + tid] = S.beta[1][tid];
  }
  if (tid == 0) {
    hard[7] = fallback;
    int* info = a->info + (size_t)prob * 8;
    info[0] = error ? error : reason;
    info[1] = converged;
    info[2] = it;
    info[3] = fallback ? feasible_it : -1;
    info[4] = (int)physical;
    info[5] = (int)logical;
    info[6] = error != 0;
    info[7] = finished;
    double* sc = a->scalars + (size_t)prob * kSplitScalars;
    sc[2] = a->adaptive ? alpha_ctrl : a->alpha0;
    sc[3] = C.T;
    sc[4] = S.packet[4];  // ridge_effective of the last soft fit, left/right
    sc[5] = S.packet[5];
  }
}

}  // namespace art

using namespace art;

extern "C" int art_split_optimize(const art_split_args* args, int problems, hipStream_t stream) {
  if (!args || problems < 1) return fail(ART_ERR_ARG, "art_split_optimize: bad arguments");
  const int p = args->basis.p;
  const int d = args->d;
  if (p < 1 || p > kSplitMaxP || d < 1 || d > kSplitMaxD)
    return fail(ART_ERR_ARG, "art_split_optimize: requires 1 <= p <= 32 and 1 <= d <= 31");
  if ((!args->row_offsets && args->n < 1) || args->H < 1 || args->max_iters < 0 ||
      args->max_backtracks < 0 ||
      args->max_failures < 1 || args->mode < 0 || args->mode > 2 || args->anneal_iters < 0)
    return fail(ART_ERR_ARG, "art_split_optimize: invalid sizes or controls");
  if (args->child_precision != ART_PRECISION_FP64)
    return fail(ART_ERR_ARG,
                "art_split_optimize: FP64 child systems only (split-bf16 children are an "
                "option of art_split_anneal_batched)");
  if (!args->X || !args->y || !args->temperature || !args->theta0 || !args->node ||
      !args->theta_out || !args->beta_out || !args->hard || !args->info || !args->history ||
      !args->scalars)
    return fail(ART_ERR_ARG, "art_split_optimize: null device pointer");
  // Gram tiles cover the non-constant features only.
  const int nf = p - (args->basis.bias ? 1 : 0);
  const dim3 grid(problems);
  const PartSync one{1, 0, 0, 0, nullptr, nullptr, nullptr, 0, 0};
  if (nf <= 16) split_optimize_kernel<1><<<grid, dim3(kThreadsT<1>), 0, stream>>>(*args, one);
  else split_optimize_kernel<2><<<grid, dim3(kThreadsT<2>), 0, stream>>>(*args, one);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "split_optimize_kernel launch");
}

// Exchange slot of one part: the longest exchanged vector (both sides'
// packed augmented systems and a flag; the statistics pass's 3 + d sums; a
// trial pass's losses), rounded up to 32 doubles.
static int part_stride(int p) {
  const int nel = (p + 1) * (p + 2) / 2;
  const int longest = std::max(2 * nel + 1, std::max(3 + kSplitMaxD, kSplitMaxH));
  return (longest + 31) / 32 * 32;
}

extern "C" int art_split_part_doubles(int p, int parts) {
  if (p < 1 || p > kSplitMaxP || parts < 1) return -1;
  return 2 * parts * part_stride(p);
}

extern "C" int art_split_max_parts(const art_split_args* args, int problems) {
  if (!args || problems < 1) return 1;
  const int p = args->basis.p;
  if (p < 1 || p > kSplitMaxP) return 1;
  const int nf = p - (args->basis.bias ? 1 : 0);  // the launch's kernel: see art_split_optimize
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 1;
  const hipError_t e =
      nf <= 16 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, split_optimize_kernel<1>,
                                                               kThreadsT<1>, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, split_optimize_kernel<2>,
                                                               kThreadsT<2>, 0);
  if (e != hipSuccess || per_cu < 1) return 1;
  const long capacity = (long)per_cu * cus;
  const long parts = capacity / problems;
  return (int)(parts < 1 ? 1 : parts > kSplitMaxParts ? kSplitMaxParts : parts);
}

extern "C" int art_split_optimize_parts(const art_split_args* args, int problems, int parts,
                                        int* count, double* buf, int* failed,
                                        hipStream_t stream) {
  if (parts == 1) return art_split_optimize(args, problems, stream);
  if (!args || problems < 1 || parts < 1 || parts > kSplitMaxParts || !count || !buf || !failed)
    return fail(ART_ERR_ARG, "art_split_optimize_parts: bad arguments");
  const int p = args->basis.p;
  if (p < 1 || p > kSplitMaxP || args->d < 1 || args->d > kSplitMaxD)
    return fail(ART_ERR_ARG, "art_split_optimize_parts: requires 1 <= p <= 32 and 1 <= d <= 31");
  if ((!args->row_offsets && args->n < 1) || args->H < 1 || args->max_iters < 0 ||
      args->max_backtracks < 0 || args->max_failures < 1 || args->mode < 0 || args->mode > 2 ||
      args->anneal_iters < 0 || args->child_precision != ART_PRECISION_FP64)
    return fail(ART_ERR_ARG, "art_split_optimize_parts: invalid sizes or controls");
  if (!args->X || !args->y || !args->temperature || !args->theta0 || !args->node ||
      !args->theta_out || !args->beta_out || !args->hard || !args->info || !args->history ||
      !args->scalars)
    return fail(ART_ERR_ARG, "art_split_optimize_parts: null device pointer");
  hipError_t me = hipMemsetAsync(count, 0, (size_t)problems * sizeof(int), stream);
  if (me == hipSuccess) me = hipMemsetAsync(failed, 0, sizeof(int), stream);
  if (me != hipSuccess) return hip_fail(me, "art_split_optimize_parts: clearing the counts");
  PartSync ps{parts, 0, 0, part_stride(p), count, buf, failed, 0, 0};
  art_split_args a = *args;
  void* params[] = {&a, &ps};
  const int nf = p - (args->basis.bias ? 1 : 0);
  // Every part of a problem waits for the others at each exchange, so the
  // whole grid must be resident at once: a cooperative launch, which the
  // runtime refuses (an error, nothing runs) when it cannot be.
  const hipError_t e =
      nf <= 16 ? hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&split_optimize_kernel<1>),
                                            dim3(problems * parts), dim3(kThreadsT<1>), params, 0,
                                            stream)
               : hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&split_optimize_kernel<2>),
                                            dim3(problems * parts), dim3(kThreadsT<2>), params, 0,
                                            stream);
  if (e == hipSuccess) return ART_OK;
  (void)hipGetLastError();  // a refused launch leaves no error for the next call to find
  return hip_fail(e, "split_optimize_kernel cooperative launch");
}
