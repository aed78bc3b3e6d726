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
#include <mutex>
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
  __device__ int slot(int c) const { return table[c].slot; }
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
  __device__ int slot(int c) const { return c; }
};

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

  const int slot = chunks.slot(c);
  const size_t base = ((size_t)side * nchunks + slot) * npairs;
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
  if (threadIdx.x == 0 && blockIdx.y == 0) flags[(size_t)side * nchunks + slot] = bad;
}

// ---------------------------------------------------------------------------
// Row staging for the leaf kernels.  A chunk's rows stream through LDS in
// blocks of rb rows, double buffered: the direct global->LDS copies
// (global_load_lds_dword, LDS slot = wave base + 4 * lane) of block i+1 are
// in flight while block i is computed, with one barrier per block.  One LDS
// buffer holds, in doubles:
//   [0, rb*d) X rows, [rb*d, rb*(d+1)) y, [rb*(d+1), rb*(d+2)) w, then 1.0, 0.0.
// Every column of the augmented design [phi | y | 0-padding] is the product
// of two buffer entries addressed as off + r * stride (stride 0 for the
// constants), so a lane builds a feature with two LDS reads and one multiply.
constexpr int kMaxStagedD = 24;

typedef __attribute__((address_space(3))) void* lds_void_ptr;

__host__ __device__ inline int stage_rows_per_block(int d) {
  return d <= 10 ? 256 : (d <= 16 ? 128 : 64);
}
__host__ __device__ inline int stage_buffer_doubles(int d) {
  return stage_rows_per_block(d) * (d + 2) + 2;
}

size_t leaf_stage_smem_bytes(int d) {
  const size_t n = 2 * (size_t)stage_buffer_doubles(d);
  return (n > 1024 ? n : 1024) * sizeof(double);
}

// The mixed Gram kernel stages rows through registers (one row per thread)
// into an FP32 table of two halves per row, [x_0..x_{d-1}, 1, y, 0] and the
// same scaled by sqrt(w) (each half d + 3 floats); it has no FP64 LDS
// buffers.
constexpr int kMixedRows = 256;
constexpr int kMixedMaxD = 16;
__host__ __device__ inline int rows32_stride(int d) { return d + 3; }
size_t mixed_stage_smem_bytes(int d) {
  const size_t f32 = (size_t)kMixedRows * 2 * rows32_stride(d) * sizeof(float);
  return f32 > 1024 * sizeof(double) ? f32 : 1024 * sizeof(double);
}

__device__ inline void copy_dwords_async(double* dst, const void* src, int n, int wave, int lane) {
  unsigned* d32 = reinterpret_cast<unsigned*>(dst);
  const unsigned* s32 = reinterpret_cast<const unsigned*>(src);
  for (int k = wave; k * 64 < n; k += 4) {
    const int idx = k * 64 + lane;
    if (idx < n)
      __builtin_amdgcn_global_load_lds((const void*)(s32 + idx), (lds_void_ptr)(d32 + k * 64), 4,
                                       0, 0);
  }
}

__device__ inline void stage_rows_async(double* buf, const double* X, const double* y,
                                        const double* w, int d, int rb, int64_t r0, int nrows) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  copy_dwords_async(buf, X + r0 * d, nrows * d * 2, wave, lane);
  copy_dwords_async(buf + rb * d, y + r0, nrows * 2, wave, lane);
  copy_dwords_async(buf + rb * (d + 1), w + r0, nrows * 2, wave, lane);
}

// Zero both buffers (stale rows past a chunk's end must stay finite) and set
// the constant slots.
__device__ inline void stage_init(double* smem, int d) {
  const int n = 2 * stage_buffer_doubles(d);
  for (int e = threadIdx.x; e < n; e += blockDim.x) smem[e] = 0.0;
  __syncthreads();
  if (threadIdx.x < 2) {
    const int one = stage_buffer_doubles(d) - 2;
    smem[one + threadIdx.x] = threadIdx.x == 0 ? 1.0 : 0.0;
    smem[stage_buffer_doubles(d) + one + threadIdx.x] = threadIdx.x == 0 ? 1.0 : 0.0;
  }
  __syncthreads();
}

// Buffer-relative address (offset, row stride) of one factor of column f.
struct Factor {
  int off, stride;
};

__device__ inline void column_factors(const Basis& B, int rb, int f, Factor& fa, Factor& fb) {
  const int d = B.d;
  const int one = rb * (d + 2), zero = one + 1;
  int a = -1, b = -1;
  if (f < B.p) {
    basis_factors(B, f, a, b);
    fa = a >= 0 ? Factor{a, d} : Factor{one, 0};
    fb = b >= 0 ? Factor{b, d} : Factor{one, 0};
  } else if (f == B.p) {
    fa = Factor{rb * d, 1};  // target column y
    fb = Factor{one, 0};
  } else {
    fa = Factor{zero, 0};
    fb = Factor{one, 0};
  }
}

__device__ inline double staged_feature(const double* buf, const Factor& fa, const Factor& fb,
                                        int r) {
  return buf[fa.off + r * fa.stride] * buf[fb.off + r * fb.stride];
}

// Floored weight of staged row r; 0 for rows past the block's end and for
// invalid weights (negative, NaN or inf), which set `bad`.
// Branch-free so that the LDS reads of consecutive rows can be batched; the
// slot of a row past the end holds stale finite data and is masked out.
__device__ inline double staged_weight(const double* buf, int d, int rb, int r, int nrows,
                                       double floor, int& bad) {
  const double v = buf[rb * (d + 1) + r];
  const bool inside = r < nrows;
  const bool ok = (v >= 0.0) & !isinf(v);
  bad |= (int)(inside & !ok);
  return (inside & ok) ? fmax(v, floor) : 0.0;
}

// ---------------------------------------------------------------------------
// FP64 Gram for leaves whose features are built from X (the plain leaf-ridge
// path and the exact fallback of the mixed path).  Lane (col, kq) of wave w
// takes row 16 s + 4 w + kq of step s and features 16 t + col; each lane
// builds its T feature values once per step and feeds all T(T+1)/2 tile
// pairs through v_mfma_f64_16x16x4.  `mask`, when given, skips chunks whose
// leaf is not flagged.
template <int T>
__global__ __launch_bounds__(256) void gram_leaf_f64_kernel(
    const double* __restrict__ X, const double* __restrict__ y, const double* __restrict__ w,
    Basis B, double floor, const LeafChunk* __restrict__ chunks, const int* __restrict__ mask,
    double* __restrict__ partials, int* __restrict__ flags) {
  constexpr int NP = T * (T + 1) / 2;
  if (mask != nullptr && !mask[chunks[blockIdx.x].leaf]) return;
  extern __shared__ double smem[];
  const int d = B.d, rb = stage_rows_per_block(d), bsz = stage_buffer_doubles(d);
  const int c = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, kq = lane >> 4;
  const int64_t begin = chunks[c].begin, end = chunks[c].end;

  Factor fa[T], fb[T];
#pragma unroll
  for (int t = 0; t < T; ++t) column_factors(B, rb, 16 * t + col, fa[t], fb[t]);
  f64x4 acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};

  stage_init(smem, d);
  int bad = 0;
  int64_t r0 = begin;
  int nrows = (int)(end - r0 < rb ? end - r0 : rb);
  if (nrows > 0) stage_rows_async(smem, X, y, w, d, rb, r0, nrows);
  for (int blk = 0; r0 < end; ++blk) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int64_t n0 = r0 + rb;
    const int nn = n0 < end ? (int)(end - n0 < rb ? end - n0 : rb) : 0;
    if (nn > 0) stage_rows_async(smem + ((blk + 1) & 1) * bsz, X, y, w, d, rb, n0, nn);
    const double* buf = smem + (blk & 1) * bsz;
    const int steps = (nrows + 15) >> 4;
    for (int s = 0; s < steps; ++s) {
      const int r = 16 * s + 4 * wave + kq;
      const double wt = staged_weight(buf, d, rb, r, nrows, floor, bad);
      double phi[T], a[T];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        phi[t] = staged_feature(buf, fa[t], fb[t], r);
        a[t] = wt * phi[t];
      }
      int q = 0;
#pragma unroll
      for (int t1 = 0; t1 < T; ++t1)
#pragma unroll
        for (int t2 = t1; t2 < T; ++t2) {
          acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t1], phi[t2], acc[q], 0, 0, 0);
          ++q;
        }
    }
    r0 = n0;
    nrows = nn;
  }

  double* red = smem;
  const size_t base = (size_t)chunks[c].slot * NP;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wave * 256 + (kq + 4 * r) * 16 + col] = acc[q][r];
    __syncthreads();
    const double v = ((red[tid] + red[256 + tid]) + red[512 + tid]) + red[768 + tid];
    partials[(base + q) * 256 + tid] = v;
  }
  bad = __syncthreads_or(bad);
  if (tid == 0) flags[chunks[c].slot] = bad;
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
// Row-per-thread systems (p <= kRowsMaxP): thread i owns row i of the packed
// lower triangle in LDS.  The Cholesky runs in Crout order -- column k is
// s_i = A[i][k] - sum_{j<k} L[i][j] L[k][j], a dot product of the thread's own
// row with row k (a broadcast read) -- and the substitutions are
// column-oriented with one barrier per step.  The code is compact and its
// loops are not unrolled over p.
constexpr int kRowsMaxP = 120;


// Reduce + regularize + factor + solve, same contract as solve_f64_kernel;
// optionally stores the factor rows (packed) and skips unmasked systems.
__global__ __launch_bounds__(128) void solve_rows_kernel(
    const SystemDesc* __restrict__ systems, int nchunks, int T, int npairs, int p, double ridge,
    const double* __restrict__ partials, const int* __restrict__ flags,
    const double* __restrict__ tpart, const int* __restrict__ mask, double* __restrict__ factor,
    double* __restrict__ beta, int* __restrict__ status, double* __restrict__ lambda) {
  extern __shared__ double M[];  // packed lower (p+1) x (p+1): G and the row h
  __shared__ double bc[2];
  __shared__ double red[2];
  __shared__ int ired[2];
  // Without a system table, system k is chunk k alone (resident systems).
  const SystemDesc S = systems ? systems[blockIdx.x]
                               : SystemDesc{(int)blockIdx.x, (int)blockIdx.x + 1,
                                            (int)blockIdx.x, 0};
  if (mask != nullptr && !mask[S.out]) return;
  const int i = threadIdx.x, nt = blockDim.x;
  const int nel = (p + 1) * (p + 2) / 2;
  for (int e = i; e < nel; e += nt) M[e] = 0.0;
  int bad = 0;
  for (int ch = S.chunk_begin + i; ch < S.chunk_end; ch += nt)
    bad |= flags[(size_t)S.side * nchunks + ch];
  __syncthreads();
  // Tile (t1, t2) element (a, b) is G[16 t1 + a][16 t2 + b] with t1 <= t2;
  // it lands in the packed lower slot (16 t2 + b, 16 t1 + a).  Loads are
  // issued in batches of 8 before their LDS accumulations.
  const int total = npairs * 256;
  for (int ch = S.chunk_begin; ch < S.chunk_end; ++ch) {
    const double* base = partials + ((size_t)S.side * nchunks + ch) * total;
    for (int e0 = i; e0 < total; e0 += 8 * nt) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        v[u] = e < total ? base[e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        if (e < total) {
          int t1, t2;
          pair_tiles(e >> 8, T, t1, t2);
          const int r = 16 * t1 + ((e >> 4) & 15), c = 16 * t2 + (e & 15);
          if (r <= c && c <= p) M[tri(c, r)] += v[u];
        }
      }
    }
  }
  __syncthreads();
  bad = block_or(bad, ired);
  // The trace that fixes lambda: the Gram diagonal, or the exact per-chunk
  // traces when the Gram itself is approximate.
  double tpiece = 0.0;
  if (tpart != nullptr) {
    for (int ch = S.chunk_begin + i; ch < S.chunk_end; ch += nt) tpiece += tpart[ch];
  } else if (i < p) {
    tpiece = M[tri(i, i)];
  }
  const double tr = block_sum(tpiece, red);
  const double lam = ridge == 0.0 ? 0.0 : ridge * (tr / p);
  int st = bad ? -1 : ((!isfinite(tr) || !isfinite(lam)) ? -2 : 0);
  double x = 0.0;
  if (st == 0) {
    if (i < p) M[tri(i, i)] += lam;
    __syncthreads();
    st = rows_cholesky(M, bc, p, i);
    if (st == 0) {
      int nf = 0;
      if (i < p)
        for (int j = 0; j <= i; ++j) nf |= !isfinite(M[tri(i, j)]);
      if (i < p) nf |= !isfinite(M[tri(p, i)]);
      if (block_or(nf, ired)) st = -2;
    }
    if (st == 0 && factor != nullptr && i < p) {
      double* row = factor + (size_t)S.out * (p * (p + 1) / 2) + tri(i, 0);
      for (int j = 0; j <= i; ++j) row[j] = M[tri(i, j)];
    }
    if (st == 0) {
      x = rows_solve(M, bc, p, i, i < p ? M[tri(p, i)] : 0.0);
      if (block_or(i < p && !isfinite(x), ired)) st = -3;
    }
  }
  if (i < p) beta[(size_t)S.out * p + i] = st == 0 ? x : __builtin_nan("");
  if (i == 0) {
    status[S.out] = st;
    lambda[S.out] = lam;
  }
}

// The same contract for p <= PMAX <= 64, one wave per system: the lanes
// reduce the chunk partials into the packed system in LDS (two chunks' loads
// in flight: the reduction is latency-bound), then load their rows into
// registers and factor and solve them there (wave_cholesky, wave_forward,
// wave_backward: no barrier per column).  Up to three tiles
// (p <= 48) the kernel is held to 128 VGPRs, four waves per SIMD: one round
// for 4096 leaves on 1024 SIMDs.  The few registers it then spills cost less
// than a second round (solve of 4096 leaves at p = 45: 110 -> 83 us).
template <int TT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TT <= 3 ? 4 : 1))) void solve_wave_kernel(
    const SystemDesc* __restrict__ systems, int nchunks, int T, int npairs, int p, double ridge,
    const double* __restrict__ partials, const int* __restrict__ flags,
    const double* __restrict__ tpart, const int* __restrict__ mask, double* __restrict__ factor,
    double* __restrict__ beta, int* __restrict__ status, double* __restrict__ lambda) {
  // Packed lower (p+1) x (p+1): G and the row h; at least 64 doubles, as the
  // factorization's column buffer reuses it.  One wave: LDS written by one
  // lane and read by another needs no barrier (a wave's LDS operations
  // complete in order).
  extern __shared__ double M[];
  const SystemDesc S = systems ? systems[blockIdx.x]
                               : SystemDesc{(int)blockIdx.x, (int)blockIdx.x + 1,
                                            (int)blockIdx.x, 0};
  if (mask != nullptr && !mask[S.out]) return;
  const int lane = threadIdx.x;
  int bad = 0;
  for (int ch = S.chunk_begin + lane; ch < S.chunk_end; ch += 64)
    bad |= flags[(size_t)S.side * nchunks + ch];
  bad = wave_or(bad);
  // Tile (t1, t2) element (a, b) is G[16 t1 + a][16 t2 + b], t1 <= t2.  Lane
  // l owns elements l + 64 v (v < 4 NP) of a chunk's NP tile pairs, so the
  // adds need no atomics.  Two chunks' loads are issued together
  // (latency-bound otherwise); the adds stay in chunk order.  The tiles cover
  // every packed element of the (p+1) x (p+1) system once.
  constexpr int PMAX = 16 * TT < 64 ? 16 * TT : 64, NP = TT * (TT + 1) / 2, NE = 4 * NP;
  constexpr int total = NP * 256;
  double acc[NE];
#pragma unroll
  for (int v = 0; v < NE; ++v) acc[v] = 0.0;
  const double* base = partials + (size_t)S.side * nchunks * total + lane;
  int ch = S.chunk_begin;
  for (; ch + 2 <= S.chunk_end; ch += 2) {
    double v0[NE], v1[NE];
#pragma unroll
    for (int v = 0; v < NE; ++v) {
      v0[v] = base[(size_t)ch * total + 64 * v];
      v1[v] = base[(size_t)(ch + 1) * total + 64 * v];
    }
#pragma unroll
    for (int v = 0; v < NE; ++v) acc[v] = (acc[v] + v0[v]) + v1[v];
  }
  if (ch < S.chunk_end) {
#pragma unroll
    for (int v = 0; v < NE; ++v) acc[v] += base[(size_t)ch * total + 64 * v];
  }
#pragma unroll
  for (int v = 0; v < NE; ++v) {
    int t1, t2;
    pair_tiles(v >> 2, TT, t1, t2);
    const int e = ((v & 3) << 6) + lane;  // element within the tile
    const int r = 16 * t1 + (e >> 4), c = 16 * t2 + (e & 15);
    if (r <= c && c <= p) M[tri(c, r)] = acc[v];
  }
  double tpiece = 0.0;
  if (tpart != nullptr) {
    for (int ch2 = S.chunk_begin + lane; ch2 < S.chunk_end; ch2 += 64) tpiece += tpart[ch2];
  } else if (lane < p) {
    tpiece = M[tri(lane, lane)];
  }
  const double tr = wave_sum(tpiece);
  const double lam = ridge == 0.0 ? 0.0 : ridge * (tr / p);
  int st = bad ? -1 : ((!isfinite(tr) || !isfinite(lam)) ? -2 : 0);
  // Each loop below tests lane against every column; a fresh opaque copy of
  // the lane index per loop keeps those tests from being shared (and held in
  // scalar registers) across the factorization.
  double a[PMAX];
  {
    const int me = opaque_lane(lane);
#pragma unroll
    for (int j = 0; j < PMAX; ++j) a[j] = (j <= me && me < p) ? M[tri(me, j)] : 0.0;
  }
  const double h = lane < p ? M[tri(p, lane)] : 0.0;
  double x = 0.0;
  if (st == 0) {
#pragma unroll
    for (int j = 0; j < PMAX; ++j)
      if (j == lane) a[j] += lam;
    double diag;
    st = wave_cholesky<PMAX>(a, p, lane, diag, M);  // M's rows are in a[] now
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
      // Store the factor rows (the caller's copy, and the LDS copy from which
      // every lane reads its column for the back substitution).
      const int me = opaque_lane(lane);
      double* row = factor != nullptr && me < p
                        ? factor + (size_t)S.out * (p * (p + 1) / 2) + tri(me, 0)
                        : nullptr;
#pragma unroll
      for (int j = 0; j < PMAX; ++j) {
        if (j <= me && me < p) {
          M[tri(me, j)] = a[j];
          if (row) row[j] = a[j];
        }
      }
      // One wave writes and reads M: LDS operations of a wave complete in
      // order; the fence keeps the compiler from moving the reads up.
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const int me2 = opaque_lane(lane);
#pragma unroll
      for (int k = 0; k < PMAX; ++k) a[k] = (k > me2 && k < p) ? M[tri(k, me2)] : 0.0;
      x = wave_backward<PMAX>(a, dinv, p, lane, z);
      if (wave_or(lane < p && !isfinite(x))) st = -3;
    }
  }
  if (lane < p) beta[(size_t)S.out * p + lane] = st == 0 ? x : __builtin_nan("");
  if (lane == 0) {
    status[S.out] = st;
    lambda[S.out] = lam;
  }
}

// Resident systems: leaf j's augmented system is the sum of its chunks'
// partial tiles (chunks goff[j] .. goff[j+1]); bad[j] ORs their weight flags.
__global__ __launch_bounds__(256) void reduce_leaf_partials_kernel(
    const double* __restrict__ partials, const int* __restrict__ flags,
    const int* __restrict__ goff, int total, const int* __restrict__ mask,
    double* __restrict__ grams, int* __restrict__ bad) {
  const int j = blockIdx.x;
  if (mask != nullptr && !mask[j]) return;  // a skipped leaf keeps its system and flag
  const int c0 = goff[j], c1 = goff[j + 1];
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    double v = 0.0;
    for (int c = c0; c < c1; ++c) v += partials[(size_t)c * total + e];
    grams[(size_t)j * total + e] = v;
  }
  if (threadIdx.x == 0) {
    int b = 0;
    for (int c = c0; c < c1; ++c) b |= flags[c];
    bad[j] = b;
  }
}

// ---------------------------------------------------------------------------
// Mixed precision leaf fits.
//
// 1. gram_leaf_mixed_kernel: the augmented Gram [Phi | y]^T W [Phi | y] by
//    split-bf16 MFMA (v = sqrt(w) phi = hi + mid + lo, six products, 32 rows
//    per v_mfma_f32_16x16x32_bf16), accumulated in FP32 per 32-row step and
//    then in FP64, plus the exact FP64 trace per chunk.  Partials have the
//    FP64 kernel's layout, so the same solve kernel factors them.
// 2. solve_rows_kernel (with factor output): beta0 = (G~ + lambda I)^-1 h~,
//    lambda from the exact trace.
// 3. residual_leaf_kernel: FP64 data-space residual Phi^T W (y - Phi beta0)
//    per chunk (a second pass over X, y, w).
// 4. correct_wave_kernel: beta1 = beta0 + (L L^T)^-1 (r - lambda beta0); a
//    leaf whose correction exceeds tau * |beta1| (or whose mixed solve
//    failed) is flagged for the exact FP64 path, which then runs only on the
//    flagged leaves' chunks.


// Split-bf16 Gram.  One thread per row loads a block of kMixedRows rows
// straight into registers (the next block's loads are issued before the
// current block's MFMA phase, so HBM latency hides behind it), accumulates
// the exact FP64 trace of the Gram (which fixes lambda) from
// |phi|^2 = bias + s1 [+ (s1^2 + s2)/2 for the quadratic basis],
// s1 = sum x_i^2, s2 = sum x_i^4, checks the weight, and writes the row to
// an FP32 table: [x, 1, y, 0] and sqrt(w_floored) [x, 1, y, 0], so that a
// feature value sqrt(w) x_a x_b is one product of a scaled and a plain entry.
// Wave w then takes the 32-row steps s = w, w+4, ...: lane (col, kq) holds
// rows 32 s + 8 kq + e (e < 8) of features 16 t + col as v = sqrt(w) phi, split
// v = hi + mid + lo in bf16, in the operand layout of
// v_mfma_f32_16x16x32_bf16; both operands of every tile pair are these v.
// A tile pair accumulates the six products down to 2^-16 (hi*lo, lo*hi,
// mid*mid, hi*mid, mid*hi, hi*hi: Gram entries to ~1e-8 relative) in FP32
// and is added into FP64 after every 32-row step.  The two-term split
// (~1e-5) sent five times as many leaves to the FP64 path.
template <int T, int DC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DC > 0 ? 3 : 2))) void gram_leaf_mixed_kernel(
    const double* __restrict__ X, const double* __restrict__ y, const double* __restrict__ w,
    Basis B, double floor, const LeafChunk* __restrict__ chunks, const int* __restrict__ mask,
    double* __restrict__ partials, double* __restrict__ tpart, int* __restrict__ flags) {
  constexpr int NP = T * (T + 1) / 2;
  constexpr int MD = DC > 0 ? DC : kMixedMaxD;
  if (mask != nullptr && !mask[chunks[blockIdx.x].leaf]) return;  // uniform: before any barrier
  extern __shared__ double smem[];
  __shared__ double tred[4];
  const int d = DC > 0 ? DC : B.d;
  const int S32 = rows32_stride(d), RS = 2 * S32;  // half and row strides
  float* rows32 = reinterpret_cast<float*>(smem);
  const int c = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, kq = lane >> 4;
  const int64_t begin = chunks[c].begin, end = chunks[c].end;
  const double bias = B.bias ? 1.0 : 0.0;
  const double quad = B.degree == 2 ? 0.5 : 0.0;

  int fa[T], fb[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int f = 16 * t + col;
    int a = -1, b2 = -1;
    if (f < B.p) {
      basis_factors(B, f, a, b2);
      fa[t] = a >= 0 ? a : d;
      fb[t] = b2 >= 0 ? b2 : d;
    } else {
      fa[t] = f == B.p ? d + 1 : d + 2;
      fb[t] = d;
    }
    fa[t] += S32;  // scaled half
  }
  f64x4 acc[NP];
  f32x4 z[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};
    z[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  }

  // Register staging of this thread's row of the current block.
  double xr[MD], yr = 0.0, wr = 0.0;
  auto load_row = [&](int64_t row) {
    if (row < end) {
#pragma unroll
      for (int i = 0; i < MD; ++i)
        if (DC > 0 || i < d) xr[i] = X[row * d + i];
      yr = y[row];
      wr = w[row];
    }
  };
  int bad = 0;
  double tr = 0.0;
  const NegPairs neg = neg_pairs();
  int64_t r0 = begin;
  load_row(r0 + tid);
  for (; r0 < end; r0 += kMixedRows) {
    const int nrows = (int)(end - r0 < kMixedRows ? end - r0 : kMixedRows);
    __syncthreads();  // the previous block's MFMA phase is done with the table
    {
      float* row = rows32 + tid * RS;
      float* srow = row + S32;
      if (tid < nrows) {
        const bool ok = (wr >= 0.0) & !isinf(wr);
        bad |= !ok;
        const double wt = ok ? fmax(wr, floor) : 0.0;
        // The scaled half only feeds FP32 values: sqrt(w) by v_sqrt_f32 (1 ulp)
        // and FP32 products; the trace stays FP64.
        const float sw = __builtin_amdgcn_sqrtf((float)wt);
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int i = 0; i < MD; ++i)
          if (DC > 0 || i < d) {
            const double x2 = xr[i] * xr[i];
            s1 += x2;
            s2 += x2 * x2;
            const float xf = (float)xr[i];
            row[i] = xf;
            srow[i] = sw * xf;
          }
        tr += wt * (bias + s1 + quad * (s1 * s1 + s2));
        const float yf = (float)yr;
        row[d] = 1.0f;
        row[d + 1] = yf;
        row[d + 2] = 0.0f;
        srow[d] = sw;
        srow[d + 1] = sw * yf;
        srow[d + 2] = 0.0f;
      } else {
        for (int i = 0; i < RS; ++i) row[i] = 0.0f;
      }
    }
    load_row(r0 + kMixedRows + tid);  // next block, in flight during the MFMA phase
    __syncthreads();
    const int steps = (nrows + 31) >> 5;
    for (int s = wave; s < steps; s += 4) {
      // This lane's eight rows 32 s + 8 kq + e are adjacent, so all of its
      // reads of a step are immediate offsets (< 8 RS dwords) from one
      // address per factor.
      const float* base = rows32 + (32 * s + 8 * kq) * RS;
      // Both operands are sqrt(w) phi (G = sum w phi phi^T is symmetric), so
      // each value is split once.  Rows 2k and 2k + 1 of the eight share one
      // packed word.
      u32x4 ahw[T], amw[T], alw[T];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const float p0 = base[fa[t] + 2 * k * RS] * base[fb[t] + 2 * k * RS];
          const float p1 = base[fa[t] + (2 * k + 1) * RS] * base[fb[t] + (2 * k + 1) * RS];
          unsigned hi, mid, lo;
          split3_pair(p0, p1, neg, hi, mid, lo);
          ahw[t][k] = hi;
          amw[t][k] = mid;
          alw[t][k] = lo;
        }
      }
      bf16x8 ah[T], am[T], al[T];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        ah[t] = __builtin_bit_cast(bf16x8, ahw[t]);
        am[t] = __builtin_bit_cast(bf16x8, amw[t]);
        al[t] = __builtin_bit_cast(bf16x8, alw[t]);
      }
      // Terms down to 2^-16: hi*lo + lo*hi + mid*mid + hi*mid + mid*hi +
      // hi*hi, small first; each pass over the tile pairs keeps NP
      // independent accumulator chains in flight.
      {
        const bf16x8* A1[6] = {ah, al, am, ah, am, ah};
        const bf16x8* B1[6] = {al, ah, am, am, ah, ah};
#pragma unroll
        for (int term = 0; term < 6; ++term) {
          int q = 0;
#pragma unroll
          for (int t1 = 0; t1 < T; ++t1)
#pragma unroll
            for (int t2 = t1; t2 < T; ++t2, ++q)
              z[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[term][t1], B1[term][t2], z[q],
                                                             0, 0, 0);
        }
      }
      // Into FP64 after every step: the FP32 sums are dead during the next
      // step's operand construction, which keeps the kernel within the
      // register budget of three waves per SIMD.
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          acc[q][r] += (double)z[q][r];
          z[q][r] = 0.f;
        }
    }
  }

  double* red = smem;
  const int slot = chunks[c].slot;
  const size_t pbase = (size_t)slot * NP;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wave * 256 + (4 * kq + r) * 16 + col] = acc[q][r];
    __syncthreads();
    const double v = ((red[tid] + red[256 + tid]) + red[512 + tid]) + red[768 + tid];
    partials[(pbase + q) * 256 + tid] = v;
  }
  tr = wave_sum(tr);
  if (lane == 0) tred[wave] = tr;
  bad = __syncthreads_or(bad);
  if (tid == 0) {
    flags[slot] = bad;
    tpart[slot] = ((tred[0] + tred[1]) + tred[2]) + tred[3];
  }
}

// r_chunk = sum over the chunk's rows of w (y - phi . beta) phi, in FP64.
// Lane (col, kq) holds features 16t+col of row 16 s + 4 w + kq; the dot
// product is reduced over the 16 lanes that share a row.
template <int T>
__global__ __launch_bounds__(256) void residual_leaf_kernel(
    const double* __restrict__ X, const double* __restrict__ y, const double* __restrict__ w,
    Basis B, double floor, const LeafChunk* __restrict__ chunks, const double* __restrict__ beta,
    const int* __restrict__ status, double* __restrict__ rpart, const int* __restrict__ mask) {
  const int c = blockIdx.x;
  const int leaf = chunks[c].leaf;
  if (status[leaf] != 0 || (mask && !mask[leaf])) return;
  extern __shared__ double smem[];
  __shared__ double rred[4][16 * T];
  const int d = B.d, p = B.p, rb = stage_rows_per_block(d), bsz = stage_buffer_doubles(d);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, kq = lane >> 4;
  const int64_t begin = chunks[c].begin, end = chunks[c].end;

  Factor fa[T], fb[T];
  double bt[T], rv[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int f = 16 * t + col;
    column_factors(B, rb, f, fa[t], fb[t]);
    bt[t] = f < p ? beta[(size_t)leaf * p + f] : 0.0;
    rv[t] = 0.0;
  }
  stage_init(smem, d);
  int unused = 0;
  int64_t r0 = begin;
  int nrows = (int)(end - r0 < rb ? end - r0 : rb);
  if (nrows > 0) stage_rows_async(smem, X, y, w, d, rb, r0, nrows);
  for (int blk = 0; r0 < end; ++blk) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int64_t n0 = r0 + rb;
    const int nn = n0 < end ? (int)(end - n0 < rb ? end - n0 : rb) : 0;
    if (nn > 0) stage_rows_async(smem + ((blk + 1) & 1) * bsz, X, y, w, d, rb, n0, nn);
    const double* buf = smem + (blk & 1) * bsz;
    const int steps = (nrows + 15) >> 4;
    for (int s = 0; s < steps; ++s) {
      const int r = 16 * s + 4 * wave + kq;
      const double wt = staged_weight(buf, d, rb, r, nrows, floor, unused);
      double phi[T];
      double dot = 0.0;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        phi[t] = staged_feature(buf, fa[t], fb[t], r);
        dot += phi[t] * bt[t];
      }
      dot += __shfl_xor(dot, 1, kWave);
      dot += __shfl_xor(dot, 2, kWave);
      dot += __shfl_xor(dot, 4, kWave);
      dot += __shfl_xor(dot, 8, kWave);
      const double e = wt * (buf[rb * d + r] - dot);
#pragma unroll
      for (int t = 0; t < T; ++t) rv[t] += e * phi[t];
    }
    r0 = n0;
    nrows = nn;
  }
#pragma unroll
  for (int t = 0; t < T; ++t) {
    rv[t] += __shfl_xor(rv[t], 16, kWave);
    rv[t] += __shfl_xor(rv[t], 32, kWave);
    if (kq == 0) rred[wave][16 * t + col] = rv[t];
  }
  __syncthreads();
  if (tid < 16 * T)
    rpart[(size_t)chunks[c].slot * 16 * T + tid] =
        ((rred[0][tid] + rred[1][tid]) + rred[2][tid]) + rred[3][tid];
}

// Quadratic-basis residual, one row per thread.  With the input dimension D
// known at compile time the row lives in registers, and
//   phi . beta = b0 + sum_i x_i (b_i + sum_{j>=i} b_ij x_j),
//   r += e [1, x_i, x_i x_j]   (e = w (y - phi . beta)),
// cost about (D + 1)(D + 2) FP64 FMAs per row, with no cross-lane traffic
// until the chunk ends.  Chunks here are up to kResidualChunkRows long.  The
// coefficients are uniform over the workgroup and come in by scalar loads
// (scalar registers), and each thread's next row is loaded while the current
// one is processed, so that a wave keeps two rows per lane in flight.
constexpr int kResidualChunkRows = 8192;

// Wave reduce-scatter of N <= 64 values per lane (entries >= N count as 0):
// afterwards lane l holds the sum over the wave of value l (63 shuffles
// instead of N full reductions).
template <int N>
__device__ inline double reduce_scatter(double (&v)[N], int lane) {
  static_assert(N <= 64, "one value per lane");
  double u[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) {  // h = 32 stage reads the N inputs directly
    const double lo = j < N ? v[j] : 0.0, hi = j + 32 < N ? v[j + 32] : 0.0;
    const bool upper = (lane & 32) != 0;
    u[j] = (upper ? hi : lo) + __shfl_xor(upper ? lo : hi, 32, kWave);
  }
#pragma unroll
  for (int h = 16; h >= 1; h >>= 1) {
    const bool upper = (lane & h) != 0;
#pragma unroll
    for (int j = 0; j < h; ++j) {
      const double send = upper ? u[j] : u[j + h];
      const double keep = upper ? u[j + h] : u[j];
      u[j] = keep + __shfl_xor(send, h, kWave);
    }
  }
  return u[0];
}

template <int D>
__global__ __launch_bounds__(256) void residual_quad_kernel(
    const double* __restrict__ X, const double* __restrict__ y, const double* __restrict__ w,
    Basis B, double floor, const LeafChunk* __restrict__ rchunks,
    const double* __restrict__ beta, const int* __restrict__ status,
    double* __restrict__ rpart, const int* __restrict__ mask) {
  constexpr int NQ = D * (D + 1) / 2;
  constexpr int NF = 1 + D + NQ;  // [bias, linear, quadratic]
  static_assert(NF <= 64, "reduce-scatter holds one value per lane");
  const int c = blockIdx.x;
  const int leaf = rchunks[c].leaf;
  if (status[leaf] != 0 || (mask && !mask[leaf])) return;
  __shared__ double red[4][64];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int bias = B.bias, p = B.p;
  // bl[i] (linear) and bl[D + k] (quadratic) follow the optional bias.
  const double* bl = beta + (size_t)leaf * p + bias;
  const double b0 = bias ? bl[-1] : 0.0;
  double acc[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) acc[k] = 0.0;
  const int64_t begin = rchunks[c].begin, end = rchunks[c].end;
  double xn[D], wn = 0.0, yn = 0.0;
  auto load_row = [&](int64_t row) {
    if (row < end) {
#pragma unroll
      for (int i = 0; i < D; ++i) xn[i] = X[row * D + i];
      wn = w[row];
      yn = y[row];
    }
  };
  load_row(begin + tid);
  for (int64_t row = begin + tid; row < end; row += 256) {
    double x[D];
#pragma unroll
    for (int i = 0; i < D; ++i) x[i] = xn[i];
    const double wv = wn, yv = yn;
    load_row(row + 256);
    const double wt = ((wv >= 0.0) & !isinf(wv)) ? fmax(wv, floor) : 0.0;
    double dot = b0;
    int k = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double u = bl[i];
#pragma unroll
      for (int j = i; j < D; ++j, ++k) u += bl[D + k] * x[j];
      dot += x[i] * u;
    }
    const double e = wt * (yv - dot);
    acc[0] += e;
    k = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const double ex = e * x[i];
      acc[1 + i] += ex;
#pragma unroll
      for (int j = i; j < D; ++j, ++k) acc[1 + D + k] += ex * x[j];
    }
  }
  const double total = reduce_scatter<NF>(acc, lane);
  red[wave][lane] = total;
  __syncthreads();
  // Slot stride p (the basis width, NF - 1 + bias): the correction kernel
  // reads rpart[chunk * p + feature].
  if (tid < NF) {
    const double v = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
    if (tid > 0 || bias) rpart[(size_t)rchunks[c].slot * p + tid - 1 + bias] = v;
  }
}

// One refinement step per leaf with the stored factor; flags leaves that
// need more.  The first step (mask == nullptr) accepts |delta| <= tau |beta|:
// with the low-precision Gram's contraction rho ~ |delta| / |beta| the error
// left is ~ tau^2 |beta|.  Later steps run on the flagged leaves (mask) and
// accept when the step contracts (rho = |delta_k| / |delta_{k-1}| < 1/2) and
// the error left, ~ rho |delta_k| / (1 - rho), is at most tau^2 |beta|.
// dnorm holds |delta|^2 of the last step per leaf.  One wave per leaf
// (p <= PMAX <= 64): the stored factor's rows, then its columns, are loaded
// into registers for the two substitutions.
template <int PMAX>
__global__ __launch_bounds__(64) void correct_wave_kernel(
    const SystemDesc* __restrict__ systems, const int* __restrict__ roff, int p, int rstride,
    const double* __restrict__ rpart,
    const double* __restrict__ factor, const double* __restrict__ lambda,
    double* __restrict__ beta, const int* __restrict__ status, int* __restrict__ need,
    double tau, const int* __restrict__ mask, double* __restrict__ dnorm) {
  const SystemDesc S = systems[blockIdx.x];
  const int leaf = S.out, lane = threadIdx.x;
  if (mask && !mask[leaf]) return;
  if (status[leaf] != 0) {
    if (lane == 0) need[leaf] = 1;
    return;
  }
  // The packed factor comes into LDS in one pass of coalesced loads (all in
  // flight together); its rows and then its columns are read from there.
  extern __shared__ double Fs[];
  {
    const int nf = p * (p + 1) / 2;
    const double* F = factor + (size_t)leaf * nf;
    constexpr int NV = (PMAX * (PMAX + 1) / 2 + 63) / 64;
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = lane + 64 * i < nf ? F[lane + 64 * i] : 0.0;
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (lane + 64 * i < nf) Fs[lane + 64 * i] = v[i];
    // One wave writes and reads Fs: its LDS operations complete in order; the
    // fence keeps the compiler from moving the reads up.
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  double L[PMAX];
#pragma unroll
  for (int j = 0; j < PMAX; ++j) L[j] = (j <= lane && lane < p) ? Fs[tri(lane, j)] : 0.0;
  const double dinv = lane < p ? 1.0 / Fs[tri(lane, lane)] : 1.0;
  double r = 0.0, b0 = 0.0;
  if (lane < p) {
    for (int ch = roff[leaf]; ch < roff[leaf + 1]; ++ch) r += rpart[(size_t)ch * rstride + lane];
    b0 = beta[(size_t)leaf * p + lane];
    r -= lambda[leaf] * b0;
  }
  const double z = wave_forward<PMAX>(L, dinv, p, lane, r);
#pragma unroll
  for (int k = 0; k < PMAX; ++k) L[k] = (k > lane && k < p) ? Fs[tri(k, lane)] : 0.0;
  const double delta = wave_backward<PMAX>(L, dinv, p, lane, z);
  const double b1 = b0 + delta;
  const double nd = wave_sum(lane < p ? delta * delta : 0.0);
  const double nb = wave_sum(lane < p ? b1 * b1 : 0.0);
  if (lane < p) beta[(size_t)leaf * p + lane] = b1;
  if (lane == 0) {
    if (!mask) {
      need[leaf] = !(nd <= tau * tau * nb);
    } else {
      const double prev = dnorm[leaf];
      const double rho = prev > 0.0 ? sqrt(nd / prev) : 1.0;
      const double left = rho < 1.0 ? rho * sqrt(nd) / (1.0 - rho) : INFINITY;
      need[leaf] = !(rho < 0.5 && left <= tau * tau * sqrt(nb));
    }
    dnorm[leaf] = nd;
  }
}

// Reciprocal 1-norm condition number of each leaf's regularized system
// A = L L^T from its stored factor (p <= PMAX <= 64), one wave per leaf:
// lane k forms column k of A (sum_j L_ij L_kj) and column k of A^-1 (L v =
// e_k, then L^T x = v), and rcond = 1 / (max_k |A e_k|_1 max_k |A^-1 e_k|_1).
// This is the quantity LAPACK dpocon estimates (art/models.py:288-296 uses
// the estimate); here it is exact up to rounding.  Failed leaves get 0.
template <int PMAX>
__global__ __launch_bounds__(64) void rcond_wave_kernel(const double* __restrict__ factor,
                                                         int p, const int* __restrict__ status,
                                                         double* __restrict__ rcond) {
  const int leaf = blockIdx.x, lane = threadIdx.x;
  if (status[leaf] != 0) {
    if (lane == 0) rcond[leaf] = 0.0;
    return;
  }
  extern __shared__ double Fs[];
  const int nf = p * (p + 1) / 2;
  const double* F = factor + (size_t)leaf * nf;
  for (int e = lane; e < nf; e += 64) Fs[e] = F[e];
  __syncthreads();
  const int k = lane;
  double colA = 0.0, colInv = 0.0;
  if (k < p) {
    // |A e_k|_1 with A_ik = sum_{j <= min(i,k)} L_ij L_kj.
    for (int i = 0; i < p; ++i) {
      const int m = i < k ? i : k;
      double a = 0.0;
      for (int j = 0; j <= m; ++j) a = fma(Fs[tri(i, j)], Fs[tri(k, j)], a);
      colA += fabs(a);
    }
    // v = L^-1 e_k (v_i = 0 for i < k), then x = L^-T v.
    double v[PMAX];
#pragma unroll
    for (int i = 0; i < PMAX; ++i) {
      double r = (i == k) ? 1.0 : 0.0;
      if (i > k && i < p) {
#pragma unroll
        for (int j = 0; j < i; ++j)
          if (j >= k) r = fma(-Fs[tri(i, j)], v[j], r);
      }
      v[i] = (i >= k && i < p) ? r / Fs[tri(i, i)] : 0.0;
    }
#pragma unroll
    for (int i = PMAX - 1; i >= 0; --i) {
      if (i < p) {
        double r = v[i];
#pragma unroll
        for (int j = i + 1; j < PMAX; ++j)
          if (j < p) r = fma(-Fs[tri(j, i)], v[j], r);
        v[i] = r / Fs[tri(i, i)];
        colInv += fabs(v[i]);
      }
    }
  }
  // Wave maxima of the column norms.
  for (int o = 32; o > 0; o >>= 1) {
    colA = fmax(colA, __shfl_xor(colA, o, kWave));
    colInv = fmax(colInv, __shfl_xor(colInv, o, kWave));
  }
  if (lane == 0) {
    const double prod = colA * colInv;
    rcond[leaf] = (prod > 0.0 && isfinite(prod)) ? 1.0 / prod : 0.0;
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
}  // namespace

namespace art {
int fail(int code, const char* msg) {
  std::snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  std::snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
  return ART_ERR_HIP;
}
}  // namespace art

namespace {

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
  // mixed-precision buffers (allocated on first use)
  double* d_factor = nullptr;
  double* d_rpart = nullptr;
  double* d_tpart = nullptr;
  int* d_need = nullptr;
  // residual chunking: Gram chunks per leaf (goff) and the coarser chunks of
  // the quadratic residual kernel (rchunks, roff)
  int* d_goff = nullptr;
  int* d_roff = nullptr;
  LeafChunk* d_rchunks = nullptr;
  int nrchunks = 0;
  // Leaf plans: one device allocation holds the tables and the Gram work
  // buffers (the pointers above that fall inside it are not freed alone).
  unsigned char* arena = nullptr;
  size_t arena_bytes = 0;
  double refine_tol = ART_MIXED_REFINE_TOL;
  int refine_steps = ART_MIXED_REFINE_STEPS;  // extra steps on flagged leaves
  double* d_dnorm = nullptr;
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
  auto release = [P](void* ptr) {
    const auto* b = static_cast<const unsigned char*>(ptr);
    if (ptr != nullptr && !(b >= P->arena && b < P->arena + P->arena_bytes)) (void)hipFree(ptr);
  };
  release(P->d_chunks);
  release(P->d_sys);
  release(P->d_partials);
  release(P->d_flags);
  release(P->d_scratch);
  if (P->h_sys_pinned) (void)hipHostFree(P->h_sys_pinned);
  if (P->sys_ready) (void)hipEventDestroy(P->sys_ready);
  for (hipEvent_t e : P->tev) (void)hipEventDestroy(e);
  release(P->d_factor);
  release(P->d_rpart);
  release(P->d_tpart);
  release(P->d_need);
  release(P->d_dnorm);
  release(P->d_goff);
  release(P->d_roff);
  release(P->d_rchunks);
  if (P->arena) (void)hipFree(P->arena);
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

// Sort a chunk table (built in leaf order, slot = index) into its launch
// order: longest first, ties in slot order.
static void launch_order(std::vector<LeafChunk>& table) {
  std::stable_sort(table.begin(), table.end(), [](const LeafChunk& a, const LeafChunk& b) {
    return a.end - a.begin > b.end - b.begin;
  });
}

// Gram work items of K leaves: leaf j's rows [offsets[j], offsets[j+1]) in
// chunks of at most chunk_rows, numbered (slot) in leaf order, then sorted
// into launch order -- largest chunk first, so that the long chunks do not
// start last and leave a tail (rows differ by up to 256x between leaves).
// sys[j] names leaf j's slots.  False when the offsets decrease.
static bool leaf_tables(const int64_t* offsets, int K, int chunk_rows,
                        std::vector<LeafChunk>& chunks, std::vector<SystemDesc>& sys) {
  chunks.clear();
  sys.assign(K, SystemDesc{});
  for (int j = 0; j < K; ++j) {
    const int64_t b = offsets[j], e = offsets[j + 1];
    if (e < b) return false;
    sys[j].chunk_begin = (int)chunks.size();
    for (int64_t r = b; r < e; r += chunk_rows)
      chunks.push_back(
          LeafChunk{r, r + chunk_rows < e ? r + chunk_rows : e, j, (int)chunks.size()});
    sys[j].chunk_end = (int)chunks.size();
    sys[j].out = j;
    sys[j].side = 0;
  }
  if (chunks.empty()) chunks.push_back(LeafChunk{0, 0, 0, 0});
  launch_order(chunks);
  return true;
}

int art_leaf_plan_create(const int64_t* offsets, int K, int d, int degree, int include_bias,
                         int bias_last, int chunk_rows, art_ridge_plan** out) {
  if (out == nullptr || offsets == nullptr || K < 1 || d < 1)
    return fail(ART_ERR_ARG, "art_leaf_plan_create: bad arguments");
  *out = nullptr;
  if ((degree != 1 && degree != 2) || (bias_last && degree != 1))
    return fail(ART_ERR_ARG, "art_leaf_plan_create: degree must be 1 or 2");
  if (chunk_rows < 16 || chunk_rows % 16) chunk_rows = ART_DEFAULT_LEAF_CHUNK_ROWS;
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
  std::vector<SystemDesc> sys;
  if (!leaf_tables(offsets, K, chunk_rows, chunks, sys)) {
    plan_free(P);
    return fail(ART_ERR_ARG, "art_leaf_plan_create: offsets must be nondecreasing");
  }
  P->nchunks = (int)chunks.size();
  P->nsys = K;
  std::vector<int> goff(K + 1), roff(K + 1);
  std::vector<LeafChunk> rchunks;
  for (int j = 0; j < K; ++j) {
    goff[j] = sys[j].chunk_begin;
    roff[j] = (int)rchunks.size();
    for (int64_t r = offsets[j]; r < offsets[j + 1]; r += kResidualChunkRows)
      rchunks.push_back(LeafChunk{r, r + kResidualChunkRows < offsets[j + 1] ? r + kResidualChunkRows
                                                                          : offsets[j + 1],
                                  j, (int)rchunks.size()});
  }
  goff[K] = sys[K - 1].chunk_end;
  roff[K] = (int)rchunks.size();
  if (rchunks.empty()) rchunks.push_back(LeafChunk{0, 0, 0, 0});
  P->nrchunks = (int)rchunks.size();
  launch_order(rchunks);
  // One allocation: the tables (uploaded by one copy), then the chunk weight
  // flags and traces, the chunk partial systems and, for systems too wide
  // for LDS, the solve scratch; every part starts on a 256-byte boundary.
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_off = (size_t)(K + 1) * sizeof(int);
  const size_t b_rch = rchunks.size() * sizeof(LeafChunk);
  const size_t b_ch = chunks.size() * sizeof(LeafChunk), b_sys = sys.size() * sizeof(SystemDesc);
  const size_t o_goff = 0, o_roff = up(o_goff + b_off), o_rch = up(o_roff + b_off);
  const size_t o_ch = up(o_rch + b_rch), o_sys = up(o_ch + b_ch), tables = up(o_sys + b_sys);
  const size_t o_flags = tables, o_tpart = up(o_flags + (size_t)P->nchunks * sizeof(int));
  const size_t o_part = up(o_tpart + (size_t)P->nchunks * sizeof(double));
  const size_t nel = (size_t)(P->p + 1) * (P->p + 2) / 2;
  const size_t b_part = (size_t)P->nchunks * P->npairs * 256 * sizeof(double);
  const bool scratch = nel * sizeof(double) > ART_SOLVE_LDS_BYTES;
  const size_t o_scr = up(o_part + b_part);
  const size_t total = scratch ? o_scr + (size_t)K * nel * sizeof(double) : o_scr;
  std::vector<unsigned char> host(tables, 0);
  std::memcpy(host.data() + o_goff, goff.data(), b_off);
  std::memcpy(host.data() + o_roff, roff.data(), b_off);
  std::memcpy(host.data() + o_rch, rchunks.data(), b_rch);
  std::memcpy(host.data() + o_ch, chunks.data(), b_ch);
  std::memcpy(host.data() + o_sys, sys.data(), b_sys);
  if (hipMalloc(&P->arena, total) != hipSuccess) {
    P->arena = nullptr;
    plan_free(P);
    return fail(ART_ERR_ALLOC, "art_leaf_plan_create: hipMalloc failed");
  }
  P->arena_bytes = total;
  unsigned char* A = P->arena;
  P->d_goff = reinterpret_cast<int*>(A + o_goff);
  P->d_roff = reinterpret_cast<int*>(A + o_roff);
  P->d_rchunks = reinterpret_cast<LeafChunk*>(A + o_rch);
  P->d_chunks = reinterpret_cast<LeafChunk*>(A + o_ch);
  P->d_sys = reinterpret_cast<SystemDesc*>(A + o_sys);
  P->d_flags = reinterpret_cast<int*>(A + o_flags);
  P->d_tpart = reinterpret_cast<double*>(A + o_tpart);
  P->d_partials = reinterpret_cast<double*>(A + o_part);
  if (scratch) P->d_scratch = reinterpret_cast<double*>(A + o_scr);
  if (hipMemcpy(A, host.data(), tables, hipMemcpyHostToDevice) != hipSuccess) {
    plan_free(P);
    return fail(ART_ERR_HIP, "art_leaf_plan_create: table upload failed");
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

// Row-per-lane solves: one wave in registers for p <= 64, the LDS form above.
static void launch_rows_solve(const SystemDesc* sys, int nsys, int nchunks, int T, int npairs,
                              int p, double ridge, const double* partials, const int* flags,
                              const double* tpart, const int* mask, double* factor,
                              double* beta, int* status, double* lambda, hipStream_t stream) {
  const size_t lds = (size_t)(p + 1) * (p + 2) / 2 * sizeof(double);
  // The wave kernel also uses the start of its LDS as a 64-double column.
  const size_t wlds = std::max(lds, (size_t)64 * sizeof(double));
#define ART_WAVE(TT)                                                                        \
  solve_wave_kernel<TT><<<dim3(nsys), dim3(64), wlds, stream>>>(                            \
      sys, nchunks, T, npairs, p, ridge, partials, flags, tpart, mask, factor, beta, status, \
      lambda)
  if (T == 1) ART_WAVE(1);
  else if (T == 2) ART_WAVE(2);
  else if (T == 3) ART_WAVE(3);
  else if (T == 4) ART_WAVE(4);
  else if (p <= 64) ART_WAVE(5);
  else
    solve_rows_kernel<<<dim3(nsys), dim3(64 * ((p + 63) / 64)), lds, stream>>>(
        sys, nchunks, T, npairs, p, ridge, partials, flags, tpart, mask, factor, beta, status,
        lambda);
#undef ART_WAVE
}

static int launch_solve(const art_ridge_plan* P, int nsys, int nchunks_stride, double ridge,
                        double* beta, int* status, double* lambda, hipStream_t stream,
                        const int* mask = nullptr, double* factor = nullptr,
                        const double* tpart = nullptr) {
  const int P1 = P->p + 1;
  const size_t bytes = (size_t)P1 * (P1 + 1) / 2 * sizeof(double);
  if (P->p <= kRowsMaxP) {
    launch_rows_solve(P->d_sys, nsys, nchunks_stride, P->T, P->npairs, P->p, ridge,
                      P->d_partials, P->d_flags, tpart, mask, factor, beta, status, lambda,
                      stream);
  } else if (bytes <= ART_SOLVE_LDS_BYTES) {
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


// Launch helpers dispatching on the compile-time tile count / padded width.
static void launch_gram_leaf_f64(const art_ridge_plan* P, const double* X, const double* y,
                                 const double* w, double floor, const int* mask,
                                 hipStream_t stream) {
  const size_t smem = leaf_stage_smem_bytes(P->basis.d);
  const dim3 grid(P->nchunks), block(256);
  switch (P->T) {
#define ART_CASE(TT)                                                                         \
  case TT:                                                                                   \
    gram_leaf_f64_kernel<TT><<<grid, block, smem, stream>>>(X, y, w, P->basis, floor,        \
                                                            P->d_chunks, mask,               \
                                                            P->d_partials, P->d_flags);      \
    break;
    ART_CASE(1) ART_CASE(2) ART_CASE(3) ART_CASE(4) ART_CASE(5) ART_CASE(6)
#undef ART_CASE
  }
}

// The split-bf16 Gram of every chunk (or of the chunks of leaves with
// mask[leaf] != 0) into the FP64 kernel's partial layout and the exact
// per-chunk traces (p <= 64, T <= 5, d <= kMixedMaxD; P->d_tpart allocated).
static void launch_gram_leaf_mixed(const art_ridge_plan* P, const double* X, const double* y,
                                   const double* w, double floor, const int* mask,
                                   hipStream_t stream) {
  const size_t msmem = mixed_stage_smem_bytes(P->basis.d);
  const dim3 grid(P->nchunks), block(256);
  if (P->T == 3 && P->basis.d == 8) {
    gram_leaf_mixed_kernel<3, 8><<<grid, block, msmem, stream>>>(
        X, y, w, P->basis, floor, P->d_chunks, mask, P->d_partials, P->d_tpart, P->d_flags);
    return;
  }
  switch (P->T) {
#define ART_CASE(TT)                                                                    \
  case TT:                                                                              \
    gram_leaf_mixed_kernel<TT, 0><<<grid, block, msmem, stream>>>(                      \
        X, y, w, P->basis, floor, P->d_chunks, mask, P->d_partials, P->d_tpart, P->d_flags); \
    break;
    ART_CASE(1) ART_CASE(2) ART_CASE(3) ART_CASE(4) ART_CASE(5)
#undef ART_CASE
  }
}

// Mixed (fp64_gram false): split-bf16 Gram, refinement, FP64 path for the
// leaves still flagged.  FP64-refined (fp64_gram true): FP64 Gram and
// Cholesky, then the same data-space refinement steps with the tighter
// acceptance kRefinedTol; a leaf still flagged keeps its refined solution.
constexpr double kRefinedTol = 1e-8;

static int leaf_fit_mixed(art_ridge_plan* P, const double* X, const double* y, const double* w,
                          double ridge, double floor, double* beta, int* status,
                          double* lambda, hipStream_t stream, bool fp64_gram = false) {
  if (P->p > 64 || P->T > 5 || P->basis.d > kMixedMaxD)
    return fail(ART_ERR_ARG, fp64_gram
                                 ? "art_leaf_ridge_fit: refined FP64 needs p <= 64 and d <= 16"
                                 : "art_leaf_ridge_fit: mixed precision needs p <= 64 and d <= 16");
  const int p = P->p, T = P->T;
  if (P->d_factor == nullptr) {
    ART_HIP(hipMalloc(&P->d_factor, (size_t)P->K * (p * (p + 1) / 2) * sizeof(double)));
    const size_t rp = std::max((size_t)P->nchunks * 16 * T, (size_t)P->nrchunks * p);
    ART_HIP(hipMalloc(&P->d_rpart, rp * sizeof(double)));
    if (P->d_tpart == nullptr)
      ART_HIP(hipMalloc(&P->d_tpart, (size_t)P->nchunks * sizeof(double)));
    ART_HIP(hipMalloc(&P->d_need, (size_t)P->K * sizeof(int)));
    ART_HIP(hipMalloc(&P->d_dnorm, (size_t)P->K * sizeof(double)));
  }
  const size_t smem = leaf_stage_smem_bytes(P->basis.d);
  const dim3 cgrid(P->nchunks), sgrid(P->nsys), block(256);
  timing_mark(P, 0, stream);
  if (fp64_gram) {
    launch_gram_leaf_f64(P, X, y, w, floor, nullptr, stream);
  } else {
    launch_gram_leaf_mixed(P, X, y, w, floor, nullptr, stream);
  }
  ART_HIP(hipGetLastError());
  timing_mark(P, 1, stream);
  int rc = launch_solve(P, P->nsys, P->nchunks, ridge, beta, status, lambda, stream, nullptr,
                        P->d_factor, fp64_gram ? nullptr : P->d_tpart);
  if (rc != ART_OK) return rc;
  const double tau = fp64_gram ? kRefinedTol : P->refine_tol;
  // Refinement: one FP64 residual pass and correction for every leaf, then up
  // to refine_steps more on the leaves still flagged.  The residual uses the
  // register kernel for quadratic bases with bias-first or no bias
  // (2 <= d <= 9), the feature-layout kernel otherwise.
  const bool quad = P->basis.degree == 2 && !P->basis.bias_last && P->basis.d >= 2 &&
                    P->basis.d <= 9;
  for (int step = 0; step <= P->refine_steps; ++step) {
    const int* mask = step == 0 ? nullptr : P->d_need;
    if (quad) {
      const dim3 rgrid(P->nrchunks);
      switch (P->basis.d) {
#define ART_CASE(DD)                                                                     \
  case DD:                                                                               \
    residual_quad_kernel<DD><<<rgrid, block, 0, stream>>>(X, y, w, P->basis, floor,       \
                                                          P->d_rchunks, beta, status,    \
                                                          P->d_rpart, mask);             \
    break;
        ART_CASE(2) ART_CASE(3) ART_CASE(4) ART_CASE(5) ART_CASE(6) ART_CASE(7) ART_CASE(8)
        ART_CASE(9)
#undef ART_CASE
      }
    } else {
      switch (T) {
#define ART_CASE(TT)                                                                         \
  case TT:                                                                                   \
    residual_leaf_kernel<TT><<<cgrid, block, smem, stream>>>(X, y, w, P->basis, floor,       \
                                                             P->d_chunks, beta, status,      \
                                                             P->d_rpart, mask);              \
    break;
        ART_CASE(1) ART_CASE(2) ART_CASE(3) ART_CASE(4) ART_CASE(5)
#undef ART_CASE
      }
    }
    ART_HIP(hipGetLastError());
    {
      const int* ro = quad ? P->d_roff : P->d_goff;
      const int rs = quad ? p : 16 * T;
#define ART_CORRECT(PM)                                                                     \
  correct_wave_kernel<PM><<<sgrid, dim3(64), (size_t)p * (p + 1) / 2 * sizeof(double),        \
                            stream>>>(                                                          \
      P->d_sys, ro, p, rs, P->d_rpart, P->d_factor, lambda, beta, status, P->d_need,        \
      tau, mask, P->d_dnorm)
      if (p <= 16) ART_CORRECT(16);
      else if (p <= 32) ART_CORRECT(32);
      else if (p <= 48) ART_CORRECT(48);
      else ART_CORRECT(64);
#undef ART_CORRECT
    }
    ART_HIP(hipGetLastError());
  }
  if (fp64_gram) {
    timing_mark(P, 2, stream);
    return ART_OK;
  }
  // Exact FP64 path for the flagged leaves only.
  launch_gram_leaf_f64(P, X, y, w, floor, P->d_need, stream);
  ART_HIP(hipGetLastError());
  rc = launch_solve(P, P->nsys, P->nchunks, ridge, beta, status, lambda, stream, P->d_need,
                    nullptr);
  timing_mark(P, 2, stream);
  return rc;
}

int art_leaf_ridge_fit(art_ridge_plan* P, const double* X, const double* y, const double* w,
                       double ridge, double weight_floor, int precision, double* beta,
                       int* status, double* lambda, hipStream_t stream) {
  if (!P || P->kind != 0) return fail(ART_ERR_ARG, "art_leaf_ridge_fit: not a leaf plan");
  if (!X || !y || !w || !beta || !status || !lambda)
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit: null pointer");
  if (!(ridge >= 0.0) || std::isinf(ridge) || !(weight_floor >= 0.0) || std::isinf(weight_floor))
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit: ridge and weight_floor must be finite >= 0");
  if (precision == ART_PRECISION_MIXED)
    return leaf_fit_mixed(P, X, y, w, ridge, weight_floor, beta, status, lambda, stream);
  if (precision == ART_PRECISION_FP64_REFINED)
    return leaf_fit_mixed(P, X, y, w, ridge, weight_floor, beta, status, lambda, stream, true);
  if (precision != ART_PRECISION_FP64)
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit: unsupported precision");
  timing_mark(P, 0, stream);
  if (P->T <= 6 && P->basis.d <= kMaxStagedD) {
    const size_t smem = leaf_stage_smem_bytes(P->basis.d);
    const dim3 grid(P->nchunks), block(256);
    switch (P->T) {
#define ART_LEAF_CASE(TT)                                                                    \
  case TT:                                                                                   \
    gram_leaf_f64_kernel<TT><<<grid, block, smem, stream>>>(X, y, w, P->basis, weight_floor, \
                                                            P->d_chunks, nullptr,            \
                                                            P->d_partials, P->d_flags);      \
    break;
      ART_LEAF_CASE(1)
      ART_LEAF_CASE(2)
      ART_LEAF_CASE(3)
      ART_LEAF_CASE(4)
      ART_LEAF_CASE(5)
      ART_LEAF_CASE(6)
#undef ART_LEAF_CASE
    }
  } else {
    PolySource src{X, y, w, P->basis, weight_floor};
    TableChunks map{P->d_chunks};
    const int groups = (P->npairs + kGramMaxQ - 1) / kGramMaxQ;
    gram_f64_kernel<<<dim3(P->nchunks, groups, 1), dim3(256), 0, stream>>>(
        src, map, P->nchunks, P->T, P->npairs, P->d_partials, P->d_flags);
  }
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

int art_leaf_gram(art_ridge_plan* P, const double* X, const double* y, const double* w,
                  double weight_floor, double* grams, int* bad, hipStream_t stream) {
  if (!P || P->kind != 0) return fail(ART_ERR_ARG, "art_leaf_gram: not a leaf plan");
  if (!X || !y || !w || !grams || !bad) return fail(ART_ERR_ARG, "art_leaf_gram: null pointer");
  if (!(weight_floor >= 0.0) || std::isinf(weight_floor))
    return fail(ART_ERR_ARG, "art_leaf_gram: weight_floor must be finite and >= 0");
  if (P->p > kRowsMaxP) return fail(ART_ERR_ARG, "art_leaf_gram: needs p <= 120");
  if (P->T <= 6 && P->basis.d <= kMaxStagedD) {
    launch_gram_leaf_f64(P, X, y, w, weight_floor, nullptr, stream);
  } else {
    PolySource src{X, y, w, P->basis, weight_floor};
    TableChunks map{P->d_chunks};
    const int groups = (P->npairs + kGramMaxQ - 1) / kGramMaxQ;
    gram_f64_kernel<<<dim3(P->nchunks, groups, 1), dim3(256), 0, stream>>>(
        src, map, P->nchunks, P->T, P->npairs, P->d_partials, P->d_flags);
  }
  ART_HIP(hipGetLastError());
  reduce_leaf_partials_kernel<<<dim3(P->K), dim3(256), 0, stream>>>(
      P->d_partials, P->d_flags, P->d_goff, P->npairs * 256, nullptr, grams, bad);
  ART_HIP(hipGetLastError());
  return ART_OK;
}

int art_leaf_gram_ex(art_ridge_plan* P, const double* X, const double* y, const double* w,
                     double weight_floor, int precision, const int* mask, double* grams,
                     int* bad, hipStream_t stream) {
  if (precision == ART_PRECISION_FP64 && mask == nullptr)
    return art_leaf_gram(P, X, y, w, weight_floor, grams, bad, stream);
  if (!P || P->kind != 0) return fail(ART_ERR_ARG, "art_leaf_gram_ex: not a leaf plan");
  if (!X || !y || !w || !grams || !bad)
    return fail(ART_ERR_ARG, "art_leaf_gram_ex: null pointer");
  if (!(weight_floor >= 0.0) || std::isinf(weight_floor))
    return fail(ART_ERR_ARG, "art_leaf_gram_ex: weight_floor must be finite and >= 0");
  if (precision == ART_PRECISION_MIXED) {
    if (P->p > 64 || P->T > 5 || P->basis.d > kMixedMaxD)
      return fail(ART_ERR_ARG, "art_leaf_gram_ex: mixed precision needs p <= 64 and d <= 16");
    if (P->d_tpart == nullptr)
      ART_HIP(hipMalloc(&P->d_tpart, (size_t)P->nchunks * sizeof(double)));
    launch_gram_leaf_mixed(P, X, y, w, weight_floor, mask, stream);
  } else if (precision == ART_PRECISION_FP64) {
    if (P->p > kRowsMaxP) return fail(ART_ERR_ARG, "art_leaf_gram_ex: needs p <= 120");
    if (P->T <= 6 && P->basis.d <= kMaxStagedD) {
      launch_gram_leaf_f64(P, X, y, w, weight_floor, mask, stream);
    } else {
      // The generic kernel has no mask: every chunk is formed and the reduce
      // below keeps only the selected leaves.
      PolySource src{X, y, w, P->basis, weight_floor};
      TableChunks map{P->d_chunks};
      const int groups = (P->npairs + kGramMaxQ - 1) / kGramMaxQ;
      gram_f64_kernel<<<dim3(P->nchunks, groups, 1), dim3(256), 0, stream>>>(
          src, map, P->nchunks, P->T, P->npairs, P->d_partials, P->d_flags);
    }
  } else {
    return fail(ART_ERR_ARG, "art_leaf_gram_ex: precision must be ART_PRECISION_FP64 or _MIXED");
  }
  ART_HIP(hipGetLastError());
  reduce_leaf_partials_kernel<<<dim3(P->K), dim3(256), 0, stream>>>(
      P->d_partials, P->d_flags, P->d_goff, P->npairs * 256, mask, grams, bad);
  ART_HIP(hipGetLastError());
  return ART_OK;
}

int art_gram_size(int p) {
  if (p < 1) return -1;
  const int T = tiles_for(p);
  return T * (T + 1) / 2 * 256;
}

int art_gram_solve(const double* grams, const int* bad, int K, art_basis basis, double ridge,
                   double* beta, int* status, double* lambda, hipStream_t stream) {
  if (K < 0 || !grams || !bad || !beta || !status || !lambda)
    return fail(ART_ERR_ARG, "art_gram_solve: bad arguments");
  const int p = basis.p;
  if (p < 1 || p > kRowsMaxP) return fail(ART_ERR_ARG, "art_gram_solve: needs 1 <= p <= 120");
  if (!(ridge >= 0.0) || std::isinf(ridge))
    return fail(ART_ERR_ARG, "art_gram_solve: ridge must be finite and >= 0");
  if (K == 0) return ART_OK;
  const int T = tiles_for(p), npairs = T * (T + 1) / 2;
  launch_rows_solve(nullptr, K, K, T, npairs, p, ridge, grams, bad, nullptr, nullptr, nullptr,
                    beta, status, lambda, stream);
  ART_HIP(hipGetLastError());
  return ART_OK;
}

int art_leaf_rcond(art_ridge_plan* P, const int* status, double* rcond, hipStream_t stream) {
  if (!P || P->kind != 0 || !status || !rcond)
    return fail(ART_ERR_ARG, "art_leaf_rcond: bad arguments");
  if (P->d_factor == nullptr)
    return fail(ART_ERR_ARG,
                "art_leaf_rcond: needs a previous mixed or refined FP64 fit (stored factors)");
  const int p = P->p;
  const size_t lds = (size_t)p * (p + 1) / 2 * sizeof(double);
#define ART_RCOND(PM) \
  rcond_wave_kernel<PM><<<dim3(P->K), dim3(64), lds, stream>>>(P->d_factor, p, status, rcond)
  if (p <= 16) ART_RCOND(16);
  else if (p <= 32) ART_RCOND(32);
  else if (p <= 48) ART_RCOND(48);
  else ART_RCOND(64);
#undef ART_RCOND
  ART_HIP(hipGetLastError());
  return ART_OK;
}

int art_leaf_fallback_count(art_ridge_plan* P, int* count) {
  if (!P || P->kind != 0 || !count) return fail(ART_ERR_ARG, "art_leaf_fallback_count: bad arguments");
  *count = 0;
  if (P->d_need == nullptr) return ART_OK;
  std::vector<int> need(P->K);
  ART_HIP(hipMemcpy(need.data(), P->d_need, P->K * sizeof(int), hipMemcpyDeviceToHost));
  for (int v : need) *count += v != 0;
  return ART_OK;
}

int art_leaf_set_refine_steps(art_ridge_plan* P, int steps) {
  if (!P || P->kind != 0) return fail(ART_ERR_ARG, "art_leaf_set_refine_steps: not a leaf plan");
  if (steps < 0 || steps > 16)
    return fail(ART_ERR_ARG, "art_leaf_set_refine_steps: steps must be in [0, 16]");
  P->refine_steps = steps;
  return ART_OK;
}

int art_leaf_set_refine_tol(art_ridge_plan* P, double tol) {
  if (!P || P->kind != 0 || !(tol > 0.0) || std::isinf(tol))
    return fail(ART_ERR_ARG, "art_leaf_set_refine_tol: tol must be finite and > 0");
  P->refine_tol = tol;
  return ART_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free leaf fits: every work buffer of art_leaf_ridge_fit_static
// (chunk table, leaf table, per-chunk partial systems and weight flags) lives
// in one static device array per GPU, so a call allocates no device memory.

static __device__ __align__(256) unsigned char scratch[ART_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (resolved once per device).  Host
// code uses only this address: the array's own name is a host-side stand-in.
int static_scratch(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "art_leaf_ridge_fit_static: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit_static: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "art_leaf_ridge_fit_static: hipGetSymbolAddress");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return ART_OK;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }
}  // namespace

extern "C" {

int art_leaf_ridge_fit_static(const int64_t* offsets, int K, art_basis basis, const double* X,
                              const double* y, const double* w, double ridge,
                              double weight_floor, double* beta, int* status, double* lambda,
                              hipStream_t stream) {
  if (offsets == nullptr || K < 1 || !X || !y || !w || !beta || !status || !lambda)
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit_static: bad arguments");
  if (basis.d < 1 || (basis.degree != 1 && basis.degree != 2) ||
      (basis.bias_last && basis.degree != 1) ||
      basis.p != basis_size(basis.d, basis.degree, basis.bias))
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit_static: inconsistent basis");
  if (!(ridge >= 0.0) || std::isinf(ridge) || !(weight_floor >= 0.0) || std::isinf(weight_floor))
    return fail(ART_ERR_ARG,
                "art_leaf_ridge_fit_static: ridge and weight_floor must be finite >= 0");
  const int p = basis.p, T = tiles_for(p), npairs = T * (T + 1) / 2;
  if (T > 6 || basis.d > kMaxStagedD)
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit_static: needs p <= 95 and d <= 24");
  std::vector<LeafChunk> chunks;
  std::vector<SystemDesc> sys;
  if (!leaf_tables(offsets, K, ART_DEFAULT_LEAF_CHUNK_ROWS, chunks, sys))
    return fail(ART_ERR_ARG, "art_leaf_ridge_fit_static: offsets must be nondecreasing");
  const int nchunks = (int)chunks.size();
  const size_t chunk_bytes = align256(chunks.size() * sizeof(LeafChunk));
  const size_t sys_bytes = align256(sys.size() * sizeof(SystemDesc));
  const size_t part_bytes = align256((size_t)nchunks * npairs * 256 * sizeof(double));
  const size_t flag_bytes = align256((size_t)nchunks * sizeof(int));
  if (chunk_bytes + sys_bytes + part_bytes + flag_bytes > ART_STATIC_SCRATCH_BYTES)
    return fail(ART_ERR_ARG,
                "art_leaf_ridge_fit_static: problem exceeds the static scratch "
                "(ART_STATIC_SCRATCH_BYTES); use a plan (art_leaf_plan_create)");
  unsigned char* base = nullptr;
  int rc = static_scratch(&base);
  if (rc != ART_OK) return rc;
  LeafChunk* d_chunks = reinterpret_cast<LeafChunk*>(base);
  SystemDesc* d_sys = reinterpret_cast<SystemDesc*>(base + chunk_bytes);
  double* d_partials = reinterpret_cast<double*>(base + chunk_bytes + sys_bytes);
  int* d_flags = reinterpret_cast<int*>(base + chunk_bytes + sys_bytes + part_bytes);
  // Pageable host sources: the copy has read them when the call returns, so
  // the tables may go out of scope; stream order keeps earlier calls' kernels
  // on this stream ahead of the overwrite.
  ART_HIP(hipMemcpyAsync(d_chunks, chunks.data(), chunks.size() * sizeof(LeafChunk),
                         hipMemcpyHostToDevice, stream));
  ART_HIP(hipMemcpyAsync(d_sys, sys.data(), sys.size() * sizeof(SystemDesc),
                         hipMemcpyHostToDevice, stream));
  // Every partial tile and weight flag is written by the Gram kernel before
  // the solve reads it: nothing needs clearing.
  const Basis B{basis.d, p, basis.degree, basis.bias ? 1 : 0, basis.bias_last ? 1 : 0};
  const size_t smem = leaf_stage_smem_bytes(B.d);
  switch (T) {
#define ART_CASE(TT)                                                                        \
  case TT:                                                                                  \
    gram_leaf_f64_kernel<TT><<<dim3(nchunks), dim3(256), smem, stream>>>(                   \
        X, y, w, B, weight_floor, d_chunks, nullptr, d_partials, d_flags);                  \
    break;
    ART_CASE(1) ART_CASE(2) ART_CASE(3) ART_CASE(4) ART_CASE(5) ART_CASE(6)
#undef ART_CASE
  }
  ART_HIP(hipGetLastError());
  launch_rows_solve(d_sys, K, nchunks, T, npairs, p, ridge, d_partials, d_flags, nullptr,
                    nullptr, nullptr, beta, status, lambda, stream);
  ART_HIP(hipGetLastError());
  return ART_OK;
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
