// Weighted Gram, trace-scaled ridge and Cholesky solve for batches of leaves.
//
// Pieces 1, 2 and 4 of the hot path (leaf features, weighted Gram Phi^T W Phi,
// ridge solve).  The reference computes the same quantities with
// materialized W*Phi, pointer-batched DGEMM and cuSOLVER potrf/potrs
// (art/_cuda_batched_cholesky.py:198-332, art/_cuda_batched_cholesky_kernels.cu);
// here the features are built in registers and contracted by MFMA, the right
// hand side rides along as an extra Gram column, and one workgroup reduces
// the per-chunk partial tiles and factors its system in LDS.
//
// Layout in HBM
//   X (N, d) f64 row-major, y (N,), w (N,) -- varlen leaves in CSR row ranges;
//   Phi (K, n, p) f64 row-major, y (K, n), pi (K, n) -- equal-size gated nodes.
//   partials [side][chunk][pair][16*16] f64: one 16x16 tile of the augmented
//   Gram [Phi | y]^T W [Phi | y] per upper tile pair (t1 <= t2).

#include "art_common.h"
#include "art_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

namespace art {

// ---------------------------------------------------------------------------
// Row sources: how a lane obtains feature f of row r and the row's weight.

// Raw inputs X with features built on the fly (varlen leaves).
struct PolySource {
  const double* X;
  const double* y;
  const double* w;
  Basis B;
  double floor;

  struct Col {
    int a, b, kind;  // kind 0: feature, 1: target y, 2: zero padding
  };
  __device__ Col column(int f) const {
    Col c{-1, -1, 2};
    if (f < B.p) {
      basis_factors(B, f, c.a, c.b);
      c.kind = 0;
    } else if (f == B.p) {
      c.kind = 1;
    }
    return c;
  }
  // Weights must be finite and nonnegative (art/models.py:239-240); a bad
  // weight flags its leaf and contributes nothing.
  __device__ double weight(int64_t r, int /*side*/, int& bad) const {
    double v = w[r];
    if (!(v >= 0.0) || isinf(v)) {
      bad = 1;
      return 0.0;
    }
    return fmax(v, floor);
  }
  __device__ double value(int64_t r, const Col& c) const {
    if (c.kind == 0) {
      const double* x = X + r * B.d;
      double v = c.a >= 0 ? x[c.a] : 1.0;
      if (c.b >= 0) v *= x[c.b];
      return v;
    }
    return c.kind == 1 ? y[r] : 0.0;
  }
};

// Materialized features with gates pi: side 0 weights max(1-pi, floor), side 1
// max(pi, floor); a gate that is non-finite or outside [0, 1] gives weight 0
// and flags its node (art/_cuda_batched_cholesky_kernels.cu:15-25).
struct PhiSource {
  const double* phi;
  const double* y;
  const double* pi;
  int p;
  double floor;

  struct Col {
    int f;
  };
  __device__ Col column(int f) const { return Col{f}; }
  __device__ double weight(int64_t r, int side, int& bad) const {
    double v = pi[r];
    if (!isfinite(v) || v < 0.0 || v > 1.0) {
      bad = 1;
      return 0.0;
    }
    return side == 0 ? fmax(1.0 - v, floor) : fmax(v, floor);
  }
  __device__ double value(int64_t r, const Col& c) const {
    if (c.f < p) return phi[r * p + c.f];
    return c.f == p ? y[r] : 0.0;
  }
};

// ---------------------------------------------------------------------------
// Chunk maps: chunk c -> row range [begin, end) and its owning system.

struct TableChunks {  // varlen leaves: a host-built table
  const LeafChunk* table;
  __device__ void get(int c, int64_t& begin, int64_t& end) const {
    begin = table[c].begin;
    end = table[c].end;
  }
};

struct EqualChunks {  // equal-size nodes: active ids, fixed chunks per node
  const int* ids;
  int64_t n;
  int per_node;
  int rows;
  __device__ void get(int c, int64_t& begin, int64_t& end) const {
    const int node = ids[c / per_node];
    const int64_t off = (int64_t)(c % per_node) * rows;
    begin = (int64_t)node * n + off;
    end = (int64_t)node * n + (off + rows < n ? off + rows : n);
  }
};

__host__ __device__ inline int pair_index(int t1, int t2, int T) {
  return t1 * T - t1 * (t1 - 1) / 2 + (t2 - t1);
}

__device__ inline void pair_tiles(int q, int T, int& t1, int& t2) {
  t1 = 0;
  while (q >= T - t1) {
    q -= T - t1;
    ++t1;
  }
  t2 = t1 + q;
}

// ---------------------------------------------------------------------------
// FP64 Gram kernel.  grid = (chunks, pair groups, sides), block = 4 waves.
// Each wave takes every 4th group of 4 rows; one v_mfma_f64_16x16x4 per tile
// pair consumes 4 rows.  Lane l supplies A[i=l%16][k=l/16] = w*phi_(16*t1+i)
// and B[k=l/16][j=l%16] = phi_(16*t2+j); the f64 accumulator holds
// D[l/16 + 4*r][l%16], r = 0..3 (measured on gfx950; the f32 and bf16 forms
// hold D[4*(l/16) + r][l%16] instead).  The 4 waves' tiles are summed in LDS in a
// fixed order, so partials are deterministic.
constexpr int kGramMaxQ = 6;

template <class Src, class Chunks>
__global__ __launch_bounds__(256) void gram_f64_kernel(Src src, Chunks chunks, int nchunks, int T,
                                                       int npairs, double* __restrict__ partials,
                                                       int* __restrict__ flags) {
  __shared__ double red[4][256];
  const int c = blockIdx.x;
  const int side = blockIdx.z;
  const int q0 = blockIdx.y * kGramMaxQ;
  const int nq = min(kGramMaxQ, npairs - q0);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int col = lane & 15;
  const int kq = lane >> 4;

  int64_t begin, end;
  chunks.get(c, begin, end);
  const int64_t span = end - begin;

  typename Src::Col ca[kGramMaxQ], cb[kGramMaxQ];
#pragma unroll
  for (int q = 0; q < kGramMaxQ; ++q) {
    int t1 = 0, t2 = 0;
    if (q < nq) pair_tiles(q0 + q, T, t1, t2);
    ca[q] = src.column(16 * t1 + col);
    cb[q] = src.column(16 * t2 + col);
  }

  f64x4 acc[kGramMaxQ];
#pragma unroll
  for (int q = 0; q < kGramMaxQ; ++q) acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};

  int bad = 0;
  for (int64_t s = 0; s < span; s += 16) {
    const int64_t k = s + wave * 4 + kq;
    const bool valid = k < span;
    const int64_t r = begin + k;
    const double wt = valid ? src.weight(r, side, bad) : 0.0;
#pragma unroll
    for (int q = 0; q < kGramMaxQ; ++q) {
      if (q < nq) {
        const double a = valid ? wt * src.value(r, ca[q]) : 0.0;
        const double b = valid ? src.value(r, cb[q]) : 0.0;
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
      }
    }
  }

  const size_t base = ((size_t)side * nchunks + c) * npairs;
#pragma unroll
  for (int q = 0; q < kGramMaxQ; ++q) {
    if (q < nq) {
#pragma unroll
      for (int r = 0; r < 4; ++r) red[wave][(kq + 4 * r) * 16 + col] = acc[q][r];
      __syncthreads();
      const int t = threadIdx.x;
      const double v = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
      partials[(base + q0 + q) * 256 + t] = v;
      __syncthreads();
    }
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0 && blockIdx.y == 0) flags[(size_t)side * nchunks + c] = bad;
}

// ---------------------------------------------------------------------------
// Reduce + regularize + factor + solve.  One wave per system.
//
// Status semantics follow art/_cuda_batched_cholesky_kernels.cu:29-105:
//   -1 flagged input, -2 non-finite trace/lambda or factor/rhs, +k the first
//   non-positive pivot (potrf info), -3 non-finite solution; NaN coefficients
//   on failure.  lambda = 0 for ridge 0, else ridge * trace(G) / p.
//
// The augmented system lives in packed lower-row storage: rows 0..p-1 hold G,
// row p holds h (then z, then beta) and y^T W y.

template <bool kLds>
__global__ __launch_bounds__(64) void solve_f64_kernel(const SystemDesc* __restrict__ systems,
                                                       int nchunks, int T, int npairs, int p,
                                                       double ridge,
                                                       const double* __restrict__ partials,
                                                       const int* __restrict__ flags,
                                                       double* __restrict__ scratch,
                                                       double* __restrict__ beta,
                                                       int* __restrict__ status,
                                                       double* __restrict__ lambda) {
  extern __shared__ double lds_mat[];
  const int s = blockIdx.x;
  const SystemDesc S = systems[s];
  const int P1 = p + 1;
  const int nel = P1 * (P1 + 1) / 2;
  double* A = kLds ? lds_mat : scratch + (size_t)s * nel;
  const int tid = threadIdx.x;

  // 1. Sum the partial tiles of this system's chunks in chunk order.
  for (int e = tid; e < nel; e += kWave) {
    int r = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while (tri(r + 1, 0) <= e) ++r;
    while (tri(r, 0) > e) --r;
    const int cc = e - tri(r, 0);
    const int q = pair_index(cc >> 4, r >> 4, T);
    const int idx = (cc & 15) * 16 + (r & 15);
    double v = 0.0;
    for (int ch = S.chunk_begin; ch < S.chunk_end; ++ch)
      v += partials[(((size_t)S.side * nchunks + ch) * npairs + q) * 256 + idx];
    A[e] = v;
  }
  int bad = 0;
  for (int ch = S.chunk_begin + tid; ch < S.chunk_end; ch += kWave)
    bad |= flags[(size_t)S.side * nchunks + ch];
  bad = wave_or(bad);
  __syncthreads();

  double tr = 0.0;
  for (int i = tid; i < p; i += kWave) tr += A[tri(i, i)];
  tr = wave_sum(tr);
  const double lam = ridge == 0.0 ? 0.0 : ridge * (tr / p);
  int st = bad ? -1 : ((!isfinite(tr) || !isfinite(lam)) ? -2 : 0);

  if (st == 0) {
    for (int i = tid; i < p; i += kWave) A[tri(i, i)] += lam;
    __syncthreads();
    // 2. Right-looking Cholesky of the leading p x p block.
    for (int k = 0; k < p; ++k) {
      const double dkk = A[tri(k, k)];
      if (!(dkk > 0.0)) {
        st = k + 1;
        break;
      }
      const double lkk = sqrt(dkk);
      __syncthreads();
      if (tid == 0) A[tri(k, k)] = lkk;
      for (int i = k + 1 + tid; i < p; i += kWave) A[tri(i, k)] /= lkk;
      __syncthreads();
      const int m = p - k - 1;
      for (int t = tid; t < m * m; t += kWave) {
        const int ii = t / m, jj = t - (t / m) * m;
        if (jj <= ii) {
          const int i = k + 1 + ii, j = k + 1 + jj;
          A[tri(i, j)] -= A[tri(i, k)] * A[tri(j, k)];
        }
      }
      __syncthreads();
    }
  }
  if (st == 0) {
    int nf = 0;
    for (int e = tid; e < tri(p, 0) + p; e += kWave) nf |= !isfinite(A[e]);
    nf = wave_or(nf);
    if (nf) st = -2;
  }
  if (st == 0) {
    // 3. Forward then back substitution on row p.
    double* z = A + tri(p, 0);
    for (int k = 0; k < p; ++k) {
      __syncthreads();
      const double zk = z[k] / A[tri(k, k)];
      __syncthreads();
      if (tid == 0) z[k] = zk;
      for (int i = k + 1 + tid; i < p; i += kWave) z[i] -= A[tri(i, k)] * zk;
    }
    for (int k = p - 1; k >= 0; --k) {
      __syncthreads();
      const double xk = z[k] / A[tri(k, k)];
      __syncthreads();
      if (tid == 0) z[k] = xk;
      for (int i = tid; i < k; i += kWave) z[i] -= A[tri(k, i)] * xk;
    }
    __syncthreads();
    int nf = 0;
    for (int i = tid; i < p; i += kWave) nf |= !isfinite(z[i]);
    nf = wave_or(nf);
    if (nf) st = -3;
  }
  const double* z = A + tri(p, 0);
  for (int i = tid; i < p; i += kWave)
    beta[(size_t)S.out * p + i] = st == 0 ? z[i] : __builtin_nan("");
  if (tid == 0) {
    status[S.out] = st;
    lambda[S.out] = lam;
  }
}

// ---------------------------------------------------------------------------
// Feature materialization and input checks.

__global__ __launch_bounds__(256) void features_kernel(const double* __restrict__ X, int64_t rows,
                                                       Basis B, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * B.p) return;
  const int64_t row = t / B.p;
  const int f = (int)(t - row * B.p);
  int a, b;
  basis_factors(B, f, a, b);
  double v = 1.0;
  if (a >= 0) v = X[row * B.d + a];
  if (b >= 0) v *= X[row * B.d + b];
  out[t] = v;
}

// flags[g] = 1 if group g (rows_per_group rows of cols values, plus its y
// entries) holds a non-finite value.  Flags must be zeroed beforehand.
__global__ __launch_bounds__(256) void check_finite_kernel(const double* __restrict__ a,
                                                           const double* __restrict__ y,
                                                           int64_t groups, int64_t rows_per_group,
                                                           int cols, int* __restrict__ flags) {
  const int64_t total = groups * rows_per_group * cols;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = t / cols;
    bool bad = !isfinite(a[t]);
    if (y != nullptr && t - row * cols == 0) bad |= !isfinite(y[row]);
    if (bad) flags[row / rows_per_group] = 1;
  }
}

}  // namespace art

// ===========================================================================
// Host side: plans and the C ABI.

using namespace art;

namespace {

thread_local char g_err[512];

int fail(int code, const char* msg) {
  std::snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  std::snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
  return ART_ERR_HIP;
}

#define ART_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

int tiles_for(int p) { return (p + 1 + 15) / 16; }

}  // namespace

struct art_ridge_plan {
  int kind;  // 0 varlen leaves, 1 equal-size gated nodes
  int K, p, T, npairs;
  int64_t n;  // equal-size rows per node
  int chunk_rows, per_node;
  Basis basis;
  int nchunks, nsys;
  LeafChunk* d_chunks = nullptr;
  SystemDesc* d_sys = nullptr;
  double* d_partials = nullptr;
  int* d_flags = nullptr;
  double* d_scratch = nullptr;
  // equal-size plans rebuild systems for each active set
  std::vector<SystemDesc> h_sys;
  SystemDesc* h_sys_pinned = nullptr;
  hipEvent_t sys_ready = nullptr;
  // optional per-fit kernel timing: 3 events per slot (before Gram, after
  // Gram, after solve), recorded on the fit's stream
  int timing_slots = 0;
  long timing_count = 0;
  std::vector<hipEvent_t> tev;
};

namespace {
void timing_mark(art_ridge_plan* P, int which, hipStream_t stream) {
  if (P->timing_slots <= 0) return;
  const int slot = (int)(P->timing_count % P->timing_slots);
  (void)hipEventRecord(P->tev[3 * slot + which], stream);
  if (which == 2) ++P->timing_count;
}
}  // namespace

extern "C" {

const char* art_last_error(void) { return g_err; }

int art_version(void) { return ART_ABI_VERSION; }

int art_feature_count(int d, int degree, int include_bias, int bias_last) {
  if (d < 1 || (degree != 1 && degree != 2) || (bias_last && degree != 1)) return -1;
  (void)bias_last;
  return basis_size(d, degree, include_bias);
}

int art_features_build(const double* X, int64_t rows, int d, int degree, int include_bias,
                       int bias_last, double* out, hipStream_t stream) {
  if (rows < 1 || d < 1 || X == nullptr || out == nullptr)
    return fail(ART_ERR_ARG, "art_features_build: empty input or null pointer");
  if ((degree != 1 && degree != 2) || (bias_last && degree != 1))
    return fail(ART_ERR_ARG, "art_features_build: degree must be 1 or 2; bias_last needs 1");
  Basis B{d, basis_size(d, degree, include_bias), degree, include_bias ? 1 : 0,
          bias_last ? 1 : 0};
  const int64_t total = rows * B.p;
  const int64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) return fail(ART_ERR_ARG, "art_features_build: too large");
  features_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(X, rows, B, out);
  ART_HIP(hipGetLastError());
  return ART_OK;
}

int art_check_finite(const double* a, const double* y, int64_t groups, int64_t rows_per_group,
                     int cols, int* flags, hipStream_t stream) {
  if (groups < 1 || rows_per_group < 1 || cols < 1 || a == nullptr || flags == nullptr)
    return fail(ART_ERR_ARG, "art_check_finite: bad arguments");
  ART_HIP(hipMemsetAsync(flags, 0, sizeof(int) * groups, stream));
  const int64_t total = groups * rows_per_group * cols;
  int64_t blocks = (total + 255) / 256;
  blocks = blocks > 8192 ? 8192 : blocks;
  check_finite_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a, y, groups,
                                                                        rows_per_group, cols,
                                                                        flags);
  ART_HIP(hipGetLastError());
  return ART_OK;
}

static void plan_free(art_ridge_plan* P) {
  if (!P) return;
  (void)hipFree(P->d_chunks);
  (void)hipFree(P->d_sys);
  (void)hipFree(P->d_partials);
  (void)hipFree(P->d_flags);
  (void)hipFree(P->d_scratch);
  if (P->h_sys_pinned) (void)hipHostFree(P->h_sys_pinned);
  if (P->sys_ready) (void)hipEventDestroy(P->sys_ready);
  for (hipEvent_t e : P->tev) (void)hipEventDestroy(e);
  delete P;
}

static int plan_alloc_common(art_ridge_plan* P, int sides) {
  const size_t ptiles = (size_t)sides * P->nchunks * P->npairs * 256;
  ART_HIP(hipMalloc(&P->d_partials, ptiles * sizeof(double)));
  ART_HIP(hipMalloc(&P->d_flags, (size_t)sides * P->nchunks * sizeof(int)));
  const int P1 = P->p + 1;
  const size_t nel = (size_t)P1 * (P1 + 1) / 2;
  if (nel * sizeof(double) > ART_SOLVE_LDS_BYTES)
    ART_HIP(hipMalloc(&P->d_scratch, (size_t)sides * P->K * nel * sizeof(double)));
  return ART_OK;
}

int art_leaf_plan_create(const int64_t* offsets, int K, int d, int degree, int include_bias,
                         int bias_last, int chunk_rows, art_ridge_plan** out) {
  if (out == nullptr || offsets == nullptr || K < 1 || d < 1)
    return fail(ART_ERR_ARG, "art_leaf_plan_create: bad arguments");
  *out = nullptr;
  if ((degree != 1 && degree != 2) || (bias_last && degree != 1))
    return fail(ART_ERR_ARG, "art_leaf_plan_create: degree must be 1 or 2");
  if (chunk_rows < 16 || chunk_rows % 16) chunk_rows = ART_DEFAULT_CHUNK_ROWS;
  art_ridge_plan* P = new (std::nothrow) art_ridge_plan();
  if (!P) return fail(ART_ERR_ALLOC, "art_leaf_plan_create: out of host memory");
  P->kind = 0;
  P->K = K;
  P->basis = Basis{d, basis_size(d, degree, include_bias), degree, include_bias ? 1 : 0,
                   bias_last ? 1 : 0};
  P->p = P->basis.p;
  P->T = tiles_for(P->p);
  P->npairs = P->T * (P->T + 1) / 2;
  P->chunk_rows = chunk_rows;
  std::vector<LeafChunk> chunks;
  std::vector<SystemDesc> sys(K);
  for (int j = 0; j < K; ++j) {
    const int64_t b = offsets[j], e = offsets[j + 1];
    if (e < b) {
      plan_free(P);
      return fail(ART_ERR_ARG, "art_leaf_plan_create: offsets must be nondecreasing");
    }
    sys[j].chunk_begin = (int)chunks.size();
    for (int64_t r = b; r < e; r += chunk_rows)
      chunks.push_back(LeafChunk{r, r + chunk_rows < e ? r + chunk_rows : e});
    sys[j].chunk_end = (int)chunks.size();
    sys[j].out = j;
    sys[j].side = 0;
  }
  if (chunks.empty()) chunks.push_back(LeafChunk{0, 0});
  P->nchunks = (int)chunks.size();
  P->nsys = K;
  int rc;
  if (hipMalloc(&P->d_chunks, chunks.size() * sizeof(LeafChunk)) != hipSuccess ||
      hipMalloc(&P->d_sys, sys.size() * sizeof(SystemDesc)) != hipSuccess) {
    plan_free(P);
    return fail(ART_ERR_HIP, "art_leaf_plan_create: hipMalloc failed");
  }
  if ((rc = plan_alloc_common(P, 1)) != ART_OK) {
    plan_free(P);
    return rc;
  }
  if (hipMemcpy(P->d_chunks, chunks.data(), chunks.size() * sizeof(LeafChunk),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(P->d_sys, sys.data(), sys.size() * sizeof(SystemDesc), hipMemcpyHostToDevice) !=
          hipSuccess) {
    plan_free(P);
    return fail(ART_ERR_HIP, "art_leaf_plan_create: upload failed");
  }
  *out = P;
  return ART_OK;
}

int art_gated_plan_create(int K, int64_t n, int p, int chunk_rows, art_ridge_plan** out) {
  if (out == nullptr || K < 1 || n < 1 || p < 1)
    return fail(ART_ERR_ARG, "art_gated_plan_create: bad arguments");
  *out = nullptr;
  if (chunk_rows < 16 || chunk_rows % 16) chunk_rows = ART_DEFAULT_CHUNK_ROWS;
  art_ridge_plan* P = new (std::nothrow) art_ridge_plan();
  if (!P) return fail(ART_ERR_ALLOC, "art_gated_plan_create: out of host memory");
  P->kind = 1;
  P->K = K;
  P->n = n;
  P->p = p;
  P->T = tiles_for(p);
  P->npairs = P->T * (P->T + 1) / 2;
  P->chunk_rows = chunk_rows;
  P->per_node = (int)((n + chunk_rows - 1) / chunk_rows);
  P->nchunks = K * P->per_node;
  P->nsys = 2 * K;
  int rc;
  if (hipMalloc(&P->d_sys, (size_t)2 * K * sizeof(SystemDesc)) != hipSuccess ||
      hipHostMalloc(&P->h_sys_pinned, (size_t)2 * K * sizeof(SystemDesc)) != hipSuccess ||
      hipEventCreateWithFlags(&P->sys_ready, hipEventDisableTiming) != hipSuccess) {
    plan_free(P);
    return fail(ART_ERR_HIP, "art_gated_plan_create: allocation failed");
  }
  if ((rc = plan_alloc_common(P, 2)) != ART_OK) {
    plan_free(P);
    return rc;
  }
  *out = P;
  return ART_OK;
}

void art_plan_destroy(art_ridge_plan* plan) { plan_free(plan); }

int art_plan_info(const art_ridge_plan* P, int* p, int* nchunks, int* tiles) {
  if (!P) return fail(ART_ERR_ARG, "art_plan_info: null plan");
  if (p) *p = P->p;
  if (nchunks) *nchunks = P->nchunks;
  if (tiles) *tiles = P->T;
  return ART_OK;
}

static int launch_solve(const art_ridge_plan* P, int nsys, int nchunks_stride, double ridge,
                        double* beta, int* status, double* lambda, hipStream_t stream) {
  const int P1 = P->p + 1;
  const size_t bytes = (size_t)P1 * (P1 + 1) / 2 * sizeof(double);
  if (bytes <= ART_SOLVE_LDS_BYTES) {
    solve_f64_kernel<true><<<dim3(nsys), dim3(64), bytes, stream>>>(
        P->d_sys, nchunks_stride, P->T, P->npairs, P->p, ridge, P->d_partials, P->d_flags,
        nullptr, beta, status, lambda);
  } else {
    solve_f64_kernel<false><<<dim3(nsys), dim3(64), 0, stream>>>(
        P->d_sys, nchunks_stride, P->T, P->npairs, P->p, ridge, P->d_partials, P->d_flags,
        P->d_scratch, beta, status, lambda);
  }
  ART_HIP(hipGetLastError());
  return ART_OK;
}

int art_leaf_ridge_fit(art_ridge_plan* P, const double* X, const double* y, const double* w,
                       double ridge, double weight_floor, int precision, double* beta,
                       int* status, double* lambda, hipStream_t stream) {
  if (!P || P->kind != 0) return fail(ART_ERR_ARG, "art_leaf_ridge_fit: not a leaf plan");
  if (!X || !y || !w || !beta || !status || !lambda)
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit: null pointer");
  if (!(ridge >= 0.0) || std::isinf(ridge) || !(weight_floor >= 0.0) || std::isinf(weight_floor))
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit: ridge and weight_floor must be finite >= 0");
  if (precision != ART_PRECISION_FP64)
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit: unsupported precision");
  PolySource src{X, y, w, P->basis, weight_floor};
  TableChunks map{P->d_chunks};
  const int groups = (P->npairs + kGramMaxQ - 1) / kGramMaxQ;
  timing_mark(P, 0, stream);
  gram_f64_kernel<<<dim3(P->nchunks, groups, 1), dim3(256), 0, stream>>>(
      src, map, P->nchunks, P->T, P->npairs, P->d_partials, P->d_flags);
  ART_HIP(hipGetLastError());
  timing_mark(P, 1, stream);
  const int rc = launch_solve(P, P->nsys, P->nchunks, ridge, beta, status, lambda, stream);
  timing_mark(P, 2, stream);
  return rc;
}

int art_gated_ridge_fit(art_ridge_plan* P, const double* phi, const double* y, const double* pi,
                        const int* ids_device, const int* ids_host, int count, double ridge,
                        double weight_floor, double* beta, int* status, double* lambda,
                        hipStream_t stream) {
  if (!P || P->kind != 1) return fail(ART_ERR_ARG, "art_gated_ridge_fit: not a gated plan");
  if (!phi || !y || !pi || !ids_device || !ids_host || !beta || !status || !lambda)
    return fail(ART_ERR_ARG, "art_gated_ridge_fit: null pointer");
  if (count < 1 || count > P->K)
    return fail(ART_ERR_ARG, "art_gated_ridge_fit: count out of range");
  if (!(ridge >= 0.0) || std::isinf(ridge) || !(weight_floor >= 0.0) || std::isinf(weight_floor))
    return fail(ART_ERR_ARG, "art_gated_ridge_fit: ridge and weight_floor must be finite >= 0");
  // Systems for this active set: side-major, chunks numbered by active slot.
  ART_HIP(hipEventSynchronize(P->sys_ready));  // staging buffer free again
  for (int side = 0; side < 2; ++side)
    for (int i = 0; i < count; ++i) {
      const int node = ids_host[i];
      if (node < 0 || node >= P->K) return fail(ART_ERR_ARG, "art_gated_ridge_fit: bad id");
      SystemDesc& S = P->h_sys_pinned[side * count + i];
      S.chunk_begin = i * P->per_node;
      S.chunk_end = (i + 1) * P->per_node;
      S.out = side * P->K + node;
      S.side = side;
    }
  ART_HIP(hipMemcpyAsync(P->d_sys, P->h_sys_pinned, (size_t)2 * count * sizeof(SystemDesc),
                         hipMemcpyHostToDevice, stream));
  ART_HIP(hipEventRecord(P->sys_ready, stream));
  PhiSource src{phi, y, pi, P->p, weight_floor};
  EqualChunks map{ids_device, P->n, P->per_node, P->chunk_rows};
  const int nchunks = count * P->per_node;
  const int groups = (P->npairs + kGramMaxQ - 1) / kGramMaxQ;
  timing_mark(P, 0, stream);
  gram_f64_kernel<<<dim3(nchunks, groups, 2), dim3(256), 0, stream>>>(
      src, map, nchunks, P->T, P->npairs, P->d_partials, P->d_flags);
  ART_HIP(hipGetLastError());
  timing_mark(P, 1, stream);
  const int rc = launch_solve(P, 2 * count, nchunks, ridge, beta, status, lambda, stream);
  timing_mark(P, 2, stream);
  return rc;
}

int art_plan_timing_enable(art_ridge_plan* P, int slots) {
  if (!P || slots < 0) return fail(ART_ERR_ARG, "art_plan_timing_enable: bad arguments");
  for (hipEvent_t e : P->tev) (void)hipEventDestroy(e);
  P->tev.assign((size_t)3 * slots, nullptr);
  for (auto& e : P->tev) ART_HIP(hipEventCreate(&e));
  P->timing_slots = slots;
  P->timing_count = 0;
  return ART_OK;
}

int art_plan_timing_read(art_ridge_plan* P, float* gram_ms, float* solve_ms, int max) {
  if (!P || max < 0) return fail(ART_ERR_ARG, "art_plan_timing_read: bad arguments");
  long n = P->timing_count < P->timing_slots ? P->timing_count : P->timing_slots;
  if (n > max) n = max;
  for (long i = 0; i < n; ++i) {
    ART_HIP(hipEventSynchronize(P->tev[3 * i + 2]));
    ART_HIP(hipEventElapsedTime(&gram_ms[i], P->tev[3 * i], P->tev[3 * i + 1]));
    ART_HIP(hipEventElapsedTime(&solve_ms[i], P->tev[3 * i + 1], P->tev[3 * i + 2]));
  }
  return (int)n;
}

}  // extern "C"
