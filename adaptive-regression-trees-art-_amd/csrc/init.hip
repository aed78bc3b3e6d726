// Random split starts on the device: w = a unit Gaussian direction, z = the
// median of X w over the node's rows (art/splitters.py:692-700 on the CPU;
// art/_cuda_batched_optimizer.cu:25-45 with cuRAND normals and cp.median on
// the reference's GPU path).
//
// Directions: FP64 Box-Muller pairs from Philox4x32-10 (Salmon et al., SC'11),
// key (seed_lo, seed_hi), counter (pair i, problem, counter_lo, counter_hi);
// the call's four words give u1 = ((w0 >> 5) 2^26 + (w1 >> 6) + 1) 2^-53 in
// (0, 1] and u2 = ((w2 >> 5) 2^26 + (w3 >> 6)) 2^-53 in [0, 1), and
// (r cos 2 pi u2, r sin 2 pi u2) with r = sqrt(-2 ln u1) are normals 2i, 2i+1.
// A direction of norm <= 1e-12 keeps its components with w_0 = 1, as the
// reference's unit_vector does.
// Median: radix select on order-preserving 64-bit keys of u = x.w (8 passes
// of 8 bits, LDS histograms, the bin found by a wave scan).  A node of at
// most kInitLdsRows rows scores each row once into LDS; a larger node
// recomputes x.w on every pass.  An even row count averages the two middle
// values, as numpy.median does: the upper one is the lower one again when
// its key repeats, else the smallest key above it (one more sweep).

#include "art_common.h"
#include "art_hip.h"

namespace art {

__device__ __forceinline__ uint4 init_philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c.x;
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}

__device__ inline uint64_t order_key(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ inline double key_value(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

constexpr int kInitThreads = 512;
constexpr int kInitLdsRows = 8192;  // 64 KB of keys

struct InitShared {
  double w[kSplitInitMaxD];
  unsigned hist[256];
  unsigned long long prefix, mask, above;
  unsigned rank, bin_count;
};

__device__ inline double row_score(const double* x, const double* w, int d) {
  double u = 0.0;
  for (int j = 0; j < d; ++j) u = fma(x[j], w[j], u);
  return u;
}

// Key of row r: from LDS when the node's keys are staged there, else x_r.w.
struct RowKeys {
  const double* X;
  const unsigned long long* lds;
  const double* w;
  int d;
  __device__ unsigned long long operator()(int r) const {
    return lds ? lds[r] : order_key(row_score(X + (size_t)r * d, w, d));
  }
};

// Radix select of the rank-th smallest key (0-based).  On return S.prefix is
// that key, S.rank its rank among the equal keys and S.bin_count how many
// keys equal it.
__device__ void select_key(const RowKeys& keys, int n, InitShared& S, unsigned rank) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    S.prefix = 0ull;
    S.mask = 0ull;
    S.rank = rank;
  }
  for (int shift = 56; shift >= 0; shift -= 8) {
    __syncthreads();
    for (int b = tid; b < 256; b += blockDim.x) S.hist[b] = 0u;
    __syncthreads();
    const unsigned long long prefix = S.prefix, mask = S.mask;
    for (int r = tid; r < n; r += blockDim.x) {
      const unsigned long long key = keys(r);
      if ((key & mask) == prefix) atomicAdd(&S.hist[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid < 64) {
      // Lane l owns bins 4l..4l+3; an inclusive scan of the lane sums finds
      // the one lane whose bins hold the rank.
      const unsigned k = S.rank;
      const unsigned h0 = S.hist[4 * tid], h1 = S.hist[4 * tid + 1], h2 = S.hist[4 * tid + 2],
                     h3 = S.hist[4 * tid + 3];
      const unsigned own = h0 + h1 + h2 + h3;
      unsigned inc = own;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const unsigned v = __shfl_up(inc, off, 64);
        if (tid >= off) inc += v;
      }
      const unsigned excl = inc - own;
      if (k >= excl && k < inc) {
        unsigned kk = k - excl, b = 4u * tid, c = h0;
        if (kk >= h0) {
          kk -= h0; ++b; c = h1;
          if (kk >= h1) {
            kk -= h1; ++b; c = h2;
            if (kk >= h2) { kk -= h2; ++b; c = h3; }
          }
        }
        S.rank = kk;
        S.bin_count = c;
        S.prefix = prefix | ((unsigned long long)b << shift);
        S.mask = mask | (255ull << shift);
      }
    }
  }
  __syncthreads();
}

// Smallest key strictly above `key` (all rows are scanned).
__device__ unsigned long long key_above(const RowKeys& keys, int n, InitShared& S,
                                        unsigned long long key) {
  const int tid = threadIdx.x;
  if (tid == 0) S.above = ~0ull;
  __syncthreads();
  unsigned long long best = ~0ull;
  for (int r = tid; r < n; r += blockDim.x) {
    const unsigned long long k = keys(r);
    if (k > key && k < best) best = k;
  }
  if (best != ~0ull) atomicMin(&S.above, best);
  __syncthreads();
  return S.above;
}

__global__ __launch_bounds__(kInitThreads) void split_init_kernel(
    const double* __restrict__ X, int n_all, const int64_t* __restrict__ row_offsets, int d,
    const int* __restrict__ node, uint64_t seed, uint64_t counter,
    const uint64_t* __restrict__ start_keys, double* __restrict__ theta) {
  __shared__ InitShared S;
  extern __shared__ unsigned long long init_keys[];
  const int q = blockIdx.x, tid = threadIdx.x;
  const int k = node[q];
  const int64_t r0 = row_offsets ? row_offsets[k] : (int64_t)k * n_all;
  const int n = row_offsets ? (int)(row_offsets[k + 1] - r0) : n_all;
  const double* Xn = X + (size_t)r0 * d;
  const int pairs = (d + 1) / 2;
  if (tid < pairs) {
    // Keyed problems draw from their own key with counter (pair, 0, 0, 0), so
    // a problem's start does not depend on its position in the launch.
    const uint64_t key = start_keys ? start_keys[q] : seed;
    const uint4 ctr = start_keys ? make_uint4((uint32_t)tid, 0u, 0u, 0u)
                           : make_uint4((uint32_t)tid, (uint32_t)q, (uint32_t)counter,
                                        (uint32_t)(counter >> 32));
    const uint4 r = init_philox(ctr, (uint32_t)key, (uint32_t)(key >> 32));
    const double u1 = ((double)(r.x >> 5) * 67108864.0 + (double)(r.y >> 6) + 1.0) * 0x1p-53;
    const double u2 = ((double)(r.z >> 5) * 67108864.0 + (double)(r.w >> 6)) * 0x1p-53;
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    S.w[2 * tid] = rad * c;
    if (2 * tid + 1 < d) S.w[2 * tid + 1] = rad * s;
  }
  __syncthreads();
  if (tid == 0) {
    double nn = 0.0;
    for (int j = 0; j < d; ++j) nn += S.w[j] * S.w[j];
    const double r = sqrt(nn);
    if (r <= 1e-12) S.w[0] = 1.0;
    else
      for (int j = 0; j < d; ++j) S.w[j] /= r;
  }
  __syncthreads();
  const bool staged = n <= kInitLdsRows;
  if (staged) {
    for (int r = tid; r < n; r += blockDim.x)
      init_keys[r] = order_key(row_score(Xn + (size_t)r * d, S.w, d));
  }
  const RowKeys keys{Xn, staged ? init_keys : nullptr, S.w, d};
  // select_key starts with a barrier, which also publishes the staged keys.
  select_key(keys, n, S, (unsigned)((n - 1) / 2));
  const unsigned long long lo = S.prefix;
  double z = key_value(lo);
  if (!(n & 1)) {
    const bool repeats = S.rank + 1 < S.bin_count;
    // Every thread reads S before key_above's first barrier and S.above is
    // only written after it.
    const unsigned long long hi = repeats ? lo : key_above(keys, n, S, lo);
    z = (z + key_value(hi)) * 0.5;
  }
  if (tid < d) theta[(size_t)q * (d + 1) + tid] = S.w[tid];
  if (tid == 0) theta[(size_t)q * (d + 1) + d] = z;
}

}  // namespace art

using namespace art;

namespace {

// Nodes of at most kInitLdsRows rows stage their keys in dynamic LDS; the
// allocation covers the largest possible node (n unknown to the host for
// row ranges).
int launch_init(const double* X, int n, const int64_t* row_offsets, int d, const int* node,
                int problems, uint64_t seed, uint64_t counter, const uint64_t* start_keys,
                double* theta, hipStream_t stream) {
  const int smem = (int)(sizeof(unsigned long long) *
                         (row_offsets ? kInitLdsRows : (n <= kInitLdsRows ? n : 0)));
  hipError_t e = hipFuncSetAttribute((const void*)split_init_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(sizeof(unsigned long long) * kInitLdsRows));
  if (e != hipSuccess) return hip_fail(e, "split_init_kernel shared memory attribute");
  split_init_kernel<<<problems, kInitThreads, smem, stream>>>(X, n, row_offsets, d, node, seed,
                                                              counter, start_keys, theta);
  e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "split_init_kernel launch");
}

}  // namespace

extern "C" int art_split_init(const double* X, int n, int d, const int* node, int problems,
                              uint64_t seed, uint64_t counter, double* theta,
                              hipStream_t stream) {
  if (!X || !node || !theta || n < 1 || d < 1 || d > kSplitInitMaxD || problems < 0)
    return fail(ART_ERR_ARG, "art_split_init: bad arguments (1 <= d <= 32, n >= 1)");
  if (problems == 0) return ART_OK;
  return launch_init(X, n, nullptr, d, node, problems, seed, counter, nullptr, theta, stream);
}

extern "C" int art_split_init_ranges(const double* X, const int64_t* row_offsets, int d,
                                     const int* node, int problems, uint64_t seed,
                                     uint64_t counter, const uint64_t* start_keys, double* theta,
                                     hipStream_t stream) {
  if (!X || !row_offsets || !node || !theta || d < 1 || d > kSplitInitMaxD || problems < 0)
    return fail(ART_ERR_ARG, "art_split_init_ranges: bad arguments (1 <= d <= 32)");
  if (problems == 0) return ART_OK;
  return launch_init(X, 0, row_offsets, d, node, problems, seed, counter, start_keys, theta,
                     stream);
}
