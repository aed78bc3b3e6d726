// Shared device helpers for the ART hot-path kernels (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "art_hip.h"

namespace art {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kWave = 64;
constexpr int kSplitInitMaxD = 32;  // widest node input of the random-start kernel

// Host-side error reporting shared by every entry point: the message is
// returned by art_last_error() on the calling thread.
int fail(int code, const char* msg);
int hip_fail(hipError_t e, const char* where);

// Feature-basis descriptor shared by every kernel that builds leaf features
// from raw inputs (the C ABI's art_basis).  Column order
// (art/_cuda_features.py:13-29):
//   degree 2: [1?, x_0..x_{d-1}, x_i*x_j for i<=j with i outermost]
//   degree 1: [1?, x_0..x_{d-1}] or, with bias_last, [x_0..x_{d-1}, 1]
using Basis = art_basis;

// Factor pair of feature column f: -1 stands for the constant 1.
__host__ __device__ inline void basis_factors(const Basis& B, int f, int& a, int& b) {
  a = -1;
  b = -1;
  if (B.bias) {
    if (B.bias_last) {
      if (f == B.d) return;
    } else {
      if (f == 0) return;
      f -= 1;
    }
  }
  if (f < B.d) {
    a = f;
    return;
  }
  int q = f - B.d;
  int i = 0;
  while (q >= B.d - i) {
    q -= B.d - i;
    ++i;
  }
  a = i;
  b = i + q;
}

__host__ __device__ inline int basis_size(int d, int degree, int bias) {
  return (degree == 2 ? d + d * (d + 1) / 2 : d) + (bias ? 1 : 0);
}

// Wave-wide sum by LDS-free cross-lane shuffles.
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline int wave_or(int v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
  return v;
}

// Value of v in lane `src` (src uniform across the wave).
__device__ inline double readlane_f64(double v, int src) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, src);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Row range of one Gram work item (a chunk of one leaf's rows).  The table is
// launched largest chunk first; `slot` is the chunk's place in leaf order,
// where its partial results go (a leaf's chunks have adjacent slots).
struct LeafChunk {
  int64_t begin;
  int64_t end;
  int leaf;
  int slot;
};

// One ridge system: its chunks [chunk_begin, chunk_end), output slot and side.
struct SystemDesc {
  int chunk_begin;
  int chunk_end;
  int out;
  int side;
};

// Index of element (i, j), j <= i, in packed lower-triangular row storage.
__host__ __device__ inline int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// Tile pairs (t1 <= t2) of a T-tile symmetric matrix, enumerated t1 outermost.
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
// Row-per-thread dense SPD systems in LDS (thread i owns row i of the packed
// lower triangle).  The Cholesky runs in Crout order -- column k is
// s_i = A[i][k] - sum_{j<k} L[i][j] L[k][j], a dot product of the thread's own
// row with row k (a broadcast read) -- and the substitutions are
// column-oriented with one barrier per step.  Every thread of the block calls
// these (they synchronize the block).

// Crout Cholesky of rows 0..p-1 of packed M; returns info (uniform).
__device__ inline int rows_cholesky(double* M, double* piv, int p, int i) {
  for (int k = 0; k < p; ++k) {
    double s = 0.0;
    if (i >= k && i < p) {
      const double* ri = M + tri(i, 0);
      const double* rk = M + tri(k, 0);
      double s0 = ri[k], s1 = 0.0, s2 = 0.0, s3 = 0.0;
      int j = 0;
      for (; j + 3 < k; j += 4) {
        s0 -= ri[j] * rk[j];
        s1 -= ri[j + 1] * rk[j + 1];
        s2 -= ri[j + 2] * rk[j + 2];
        s3 -= ri[j + 3] * rk[j + 3];
      }
      for (; j < k; ++j) s0 -= ri[j] * rk[j];
      s = (s0 + s1) + (s2 + s3);
    }
    if (i == k) *piv = s;
    __syncthreads();
    const double dkk = *piv;
    if (!(dkk > 0.0)) return k + 1;
    const double lkk = sqrt(dkk);
    if (i == k) M[tri(k, k)] = lkk;
    else if (i > k && i < p) M[tri(i, k)] = s / lkk;
    __syncthreads();
  }
  return 0;
}

// Solve L L^T x = b with L in packed M; b and the result are thread i's entry.
__device__ inline double rows_solve(const double* M, double* bc, int p, int i, double b) {
  double z = i < p ? b : 0.0;
  for (int k = 0; k < p; ++k) {
    if (i == k) {
      z = z / M[tri(k, k)];
      bc[k & 1] = z;
    }
    __syncthreads();
    const double zk = bc[k & 1];
    if (i > k && i < p) z -= M[tri(i, k)] * zk;
  }
  __syncthreads();
  double x = z;
  for (int k = p - 1; k >= 0; --k) {
    if (i == k) {
      x = x / M[tri(k, k)];
      bc[k & 1] = x;
    }
    __syncthreads();
    const double xk = bc[k & 1];
    if (i < k) x -= M[tri(k, i)] * xk;
  }
  return x;
}

// ---------------------------------------------------------------------------
// Wave-resident SPD systems (p <= PMAX <= 64): lane i holds row i of the lower
// triangle in registers, a[j] = A[i][j] for j <= i, and zeros above the
// diagonal and in rows i >= p.  Every loop is unrolled over PMAX, so the row
// stays in registers; values move between lanes by v_readlane (a pivot or a
// solved unknown) or a 64-double LDS column, and no barrier is needed.  All
// 64 lanes must call these together.
//
// No lane masks its updates: the factorization updates the entries above the
// diagonal too (they never feed the lower triangle) and zeroes them at the
// end, and the substitutions then include them as zeros.  Masks per column
// would be common subexpressions across the whole factorization and would
// hold 2p scalar registers at once.

// ---------------------------------------------------------------------------
// Split-bf16 operands of the low-precision MFMA products (leaf Grams and
// the annealer's child systems).

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Round two floats to bf16 (RNE) with one v_cvt_pk_bf16_f32.  The operands
// are FP32 values (products of FP32 table entries), so no double->bf16
// conversion (emulated by a round-to-odd sequence) can form; no asm barrier
// is used, since a volatile one orders every LDS read behind the previous
// pair's conversions.
__device__ inline unsigned pack_bf16(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}

// x minus the low (c = (-1, 0)) or high (c = (0, -1)) bf16 of the packed
// word u, in one v_dot2_f32_bf16: both products are exact and the remainders
// taken here are exactly representable in FP32, so the result is exact.
__device__ inline float sub_bf16(unsigned u, unsigned c, float x) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, u),
                                         __builtin_bit_cast(bf16x2, c), x, false);
}

// The bf16 pairs (-1, 0) and (0, -1) for sub_bf16, held in scalar
// registers: this compiler encodes the (-1, 0) pair as the inline constant
// -1.0, which v_dot2c_f32_bf16 reads as the FP32 bit pattern 0xbf800000, i.e.
// the pair (0, -1).  The asm (not volatile, so free to be hoisted) hides the
// value from constant folding.
struct NegPairs {
  unsigned low, high;
};
__device__ inline NegPairs neg_pairs() {
  unsigned lo = 0x0000bf80u, hi = 0xbf800000u;
  asm("" : "+s"(lo), "+s"(hi));
  return NegPairs{lo, hi};
}

// Split a pair of floats into packed bf16 hi, mid and lo words:
// v = hi + mid + lo + O(2^-27 v).  Each remainder v - hi (then r - mid) is
// exact in FP32: it has at most 16 significant bits.
__device__ inline void split3_pair(float a, float b, const NegPairs& c, unsigned& hi,
                                   unsigned& mid, unsigned& lo) {
  hi = pack_bf16(a, b);
  const float ra = sub_bf16(hi, c.low, a), rb = sub_bf16(hi, c.high, b);
  mid = pack_bf16(ra, rb);
  lo = pack_bf16(sub_bf16(mid, c.low, ra), sub_bf16(mid, c.high, rb));
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Device (agent) scope store and load of a double: global_store /
// global_load with the device-scope cache policy, so that a workgroup on
// another XCD reads the value once the store has completed (each XCD has its
// own L2; no L2 write-back or invalidate is needed).
__device__ inline void store_dev(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double load_dev(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lane `src`'s value of v (src uniform).
__device__ inline double lane_value(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned)lo));
}

// A copy of the lane index the compiler cannot relate to other copies, so
// that `lane == k` tests of different phases are not merged (and kept alive)
// across them.  A register move only.
__device__ inline int opaque_lane(int lane) {
  int r;
  asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(lane));
  return r;
}

// Right-looking Cholesky, column by column: the pivot comes from lane k, each
// lane scales its own entry of column k and stores it in `col` (64 doubles of
// LDS), and every lane then reads the column back (broadcast reads) for its
// trailing update a[j] -= l_ik l_jk.  Sets info to k + 1 for the first pivot
// <= 0 (uniform), as rows_cholesky does, and diag to the lane's own l_ii.
// Column K is its own instantiation so that every loop has a constant trip
// count and unrolls, keeping a[] in registers (one loop nest over all columns
// exceeds the unroller's budget).  Columns j >= p are updated too, with
// zeros: a branch per column would cost more than the extra multiply-adds.
template <int PMAX, int K>
__device__ __attribute__((always_inline)) inline void wave_cholesky_col(double (&a)[PMAX], int p,
                                                                        int me, int& info,
                                                                        double& diag,
                                                                        double* col) {
  if constexpr (K < PMAX) {
    if (K < p && info == 0) {
      const double dkk = lane_value(a[K], K);
      if (!(dkk > 0.0)) {
        info = K + 1;
      } else {
        const double lkk = sqrt(dkk);
        const double lik = me == K ? lkk : a[K] / lkk;
        diag = me == K ? lkk : diag;
        a[K] = lik;
        // One wave writes and reads col: its LDS operations complete in
        // order, and the fences keep the compiler from reordering them.
        col[me] = lik;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int j = K + 1; j < PMAX; ++j) a[j] = fma(-lik, col[j], a[j]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
    }
    wave_cholesky_col<PMAX, K + 1>(a, p, me, info, diag, col);
  }
}

template <int PMAX>
__device__ __attribute__((always_inline)) inline int wave_cholesky(double (&a)[PMAX], int p,
                                                                   int lane, double& diag,
                                                                   double* col) {
  int info = 0;
  diag = 1.0;
  wave_cholesky_col<PMAX, 0>(a, p, opaque_lane(lane), info, diag, col);
  const int me = opaque_lane(lane);
#pragma unroll
  for (int j = 0; j < PMAX; ++j) a[j] = j <= me ? a[j] : 0.0;
  return info;
}

// Forward substitution L z = b: lane i holds row i of L (zeros above the
// diagonal and in rows i >= p) and b_i; dinv = 1 / l_ii (1 for i >= p).  z_k is broadcast from lane k and
// lanes below it subtract l_ik z_k (lanes above hold l_ik = 0).
template <int PMAX>
__device__ __attribute__((always_inline)) inline double wave_forward(const double (&row)[PMAX],
                                                                     double dinv, int p, int lane,
                                                                     double b) {
  const int me = opaque_lane(lane);
  // Steps k >= p change nothing (z_k = 0 and row[k] = 0 there), so they run
  // unguarded: straight-line code.
  double z = lane < p ? b : 0.0;
#pragma unroll
  for (int k = 0; k < PMAX; ++k) {
    const double zk = lane_value(z * dinv, k);  // lane k scales, all read it
    z = me == k ? zk : fma(-row[k], zk, z);
  }
  return z;
}

// Back substitution L^T x = z: lane i holds column i of L below the diagonal,
// col[k] = l_ki for k > i and 0 otherwise.  x_k is broadcast from lane k and
// lanes above it subtract l_ki x_k.
template <int PMAX>
__device__ __attribute__((always_inline)) inline double wave_backward(const double (&col)[PMAX],
                                                                      double dinv, int p, int lane,
                                                                      double z) {
  const int me = opaque_lane(lane);
  double x = z;
#pragma unroll
  for (int k = PMAX - 1; k >= 0; --k) {
    const double xk = lane_value(x * dinv, k);
    x = me == k ? xk : fma(-col[k], xk, x);
  }
  return x;
}

__device__ inline int block_or(int v, int* scratch) {
  v = wave_or(v);
  if (blockDim.x == 64) return v;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r |= scratch[w];
  __syncthreads();
  return r;
}

__device__ inline double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  if (blockDim.x == 64) return v;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += scratch[w];
  __syncthreads();
  return r;
}

}  // namespace art
