// Row movement and node statistics of the serving loop's data-driven tree fits.
//
// A fit keeps every frontier node's rows contiguous in one level buffer per
// basis group (X (N, d), y (N,), w (N,) row-major float64).  Per level:
//
// art_gram_sse        weighted SSE and weight sum of each node's ridge fit,
//                     read from its resident augmented system
//                     [Phi | y]^T W [Phi | y] (tile layout of art_leaf_gram):
//                     SSE = y'Wy - 2 beta'h + beta'G beta.
// art_gather_strided  the split search's row sample: row i of node k's
//                     sample is row floor(i n_k / m_k) of the node.
// art_partition_rows  stable partition of each split node's rows by its
//                     hyperplane (right when w.x - z >= 0, the routing rule of
//                     art/tree.py:45-50 and art_tree_leaf), left child first,
//                     into the next level buffer.  Three kernels: side flags
//                     and right counts per 1024-row block, an exclusive scan
//                     of the block counts per node, and the scatter, which
//                     ranks rows inside a block by wave ballots.
//
// Every node's result depends on its own rows only (no cross-node atomics),
// so a fit gives the same tree whatever else shares the launches.

#include "art_common.h"
#include "art_hip.h"

namespace art {

namespace {

constexpr int kPartThreads = 256;
constexpr int kPartRows = 1024;  // rows per partition block (4 rounds of 256)

__global__ __launch_bounds__(256) void gram_sse_kernel(const double* __restrict__ grams, int T,
                                                       int npairs, int p, int bias_col,
                                                       const double* __restrict__ beta,
                                                       double* __restrict__ sse,
                                                       double* __restrict__ wsum) {
  __shared__ double red[4];
  const int k = blockIdx.x, tid = threadIdx.x;
  const double* G = grams + (size_t)k * npairs * 256;
  const double* b = beta + (size_t)k * p;
  double acc = 0.0;
  for (int e = tid; e < npairs * 256; e += blockDim.x) {
    int t1, t2;
    pair_tiles(e >> 8, T, t1, t2);
    const int r = 16 * t1 + ((e >> 4) & 15), c = 16 * t2 + (e & 15);
    if (r > c || c > p) continue;
    const double g = G[e];
    if (c < p)
      acc += (r == c ? 1.0 : 2.0) * b[r] * g * b[c];
    else if (r < p)
      acc -= 2.0 * b[r] * g;
    else
      acc += g;
  }
  acc = wave_sum(acc);
  if ((tid & 63) == 0) red[tid >> 6] = acc;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
    sse[k] = s;
    if (wsum) {
      double ws = 0.0;
      if (bias_col >= 0) {
        const int t = bias_col >> 4, a = bias_col & 15;
        ws = G[(size_t)pair_index(t, t, T) * 256 + a * 16 + a];
      }
      wsum[k] = ws;
    }
  }
}

// Index of the segment of offsets[0..K] (nondecreasing) that holds i.
__device__ inline int segment_of(const int64_t* __restrict__ offsets, int K, int64_t i) {
  int lo = 0, hi = K - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (offsets[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void gather_strided_kernel(
    const double* __restrict__ X, const double* __restrict__ y, const double* __restrict__ w,
    int d, int K, const int64_t* __restrict__ src_begin, const int64_t* __restrict__ src_end,
    const int64_t* __restrict__ dst, int64_t total, double* __restrict__ Xo,
    double* __restrict__ yo, double* __restrict__ wo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int k = segment_of(dst, K, i);
  const int64_t m = dst[k + 1] - dst[k], n = src_end[k] - src_begin[k];
  const int64_t r = src_begin[k] + ((i - dst[k]) * n) / m;
  for (int j = 0; j < d; ++j) Xo[i * d + j] = X[r * d + j];
  yo[i] = y[r];
  if (wo) wo[i] = w[r];
}

// Score of one row against a node's hyperplane, summed as art_tree_leaf does.
__device__ inline bool goes_right(const double* __restrict__ x, const double* __restrict__ nw,
                                  double z, int d) {
  double s = 0.0;
  for (int j = 0; j < d; ++j) s = fma(nw[j], x[j], s);
  return (s - z) >= 0.0;
}

struct PartArgs {
  const double* X;
  const double* y;
  const double* w;
  int d, K;
  const int64_t* src_begin;  // (K) source row ranges [src_begin[k], src_end[k])
  const int64_t* src_end;
  const int64_t* blk;   // (K+1) block ranges: node k owns blocks blk[k] .. blk[k+1]
  const int64_t* dst;   // (K) destination base of node k's children
  const double* node_w; // (K, d)
  const double* node_z; // (K)
  unsigned char* flags; // (rows of the source buffer) side of each row
  int* blk_right;       // (blocks) right rows per block, then their exclusive scan
  int64_t* node_right;  // (K) right rows per node
  double* Xo;
  double* yo;
  double* wo;
};

__global__ __launch_bounds__(kPartThreads) void partition_count_kernel(PartArgs a) {
  __shared__ int red[kPartThreads / 64];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int k = segment_of(a.blk, a.K, b);
  const int64_t r0 = a.src_begin[k] + (int64_t)(b - a.blk[k]) * kPartRows;
  const int64_t r1 = min(r0 + kPartRows, a.src_end[k]);
  const double* nw = a.node_w + (size_t)k * a.d;
  const double z = a.node_z[k];
  int right = 0;
  for (int64_t r = r0 + tid; r < r1; r += kPartThreads) {
    const bool rt = goes_right(a.X + r * a.d, nw, z, a.d);
    a.flags[r] = rt;
    right += rt;
  }
  for (int off = 32; off > 0; off >>= 1) right += __shfl_down(right, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = right;
  __syncthreads();
  if (tid == 0) a.blk_right[b] = red[0] + red[1] + red[2] + red[3];
}

// One thread per node: exclusive scan of its blocks' right counts in place.
__global__ __launch_bounds__(256) void partition_scan_kernel(PartArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.K) return;
  int64_t run = 0;
  for (int64_t b = a.blk[k]; b < a.blk[k + 1]; ++b) {
    const int c = a.blk_right[b];
    a.blk_right[b] = (int)run;
    run += c;
  }
  a.node_right[k] = run;
}

__global__ __launch_bounds__(kPartThreads) void partition_scatter_kernel(PartArgs a) {
  __shared__ int wave_right[kPartThreads / 64];
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = segment_of(a.blk, a.K, b);
  const int64_t local0 = (int64_t)(b - a.blk[k]) * kPartRows;
  const int64_t r0 = a.src_begin[k] + local0;
  const int64_t r1 = min(r0 + kPartRows, a.src_end[k]);
  const int64_t n = a.src_end[k] - a.src_begin[k];
  const int64_t n_left = n - a.node_right[k];
  int64_t right_before = a.blk_right[b];          // right rows of the node before this round
  int64_t left_before = local0 - right_before;    // left rows of the node before this round
  const int d = a.d;
  for (int64_t base = r0; base < r1; base += kPartThreads) {
    const int64_t r = base + tid;
    const bool valid = r < r1;
    const bool rt = valid && a.flags[r];
    const uint64_t mask = __ballot(rt);
    const uint64_t below = lane ? (mask << (64 - lane)) : 0ull;
    const int rank_in_wave = __popcll(below);
    if (lane == 0) wave_right[wave] = __popcll(mask);
    __syncthreads();
    int rank = rank_in_wave, round_right = 0;
    for (int q = 0; q < kPartThreads / 64; ++q) {
      const int c = wave_right[q];
      if (q < wave) rank += c;
      round_right += c;
    }
    if (valid) {
      const int64_t pos = tid;  // position in this round
      const int64_t dest = rt ? a.dst[k] + n_left + right_before + rank
                              : a.dst[k] + left_before + (pos - rank);
      for (int j = 0; j < d; ++j) a.Xo[dest * d + j] = a.X[r * d + j];
      a.yo[dest] = a.y[r];
      if (a.wo) a.wo[dest] = a.w[r];
    }
    const int64_t round_rows = min((int64_t)kPartThreads, r1 - base);
    right_before += round_right;
    left_before += round_rows - round_right;
    __syncthreads();
  }
}

}  // namespace

}  // namespace art

using namespace art;

extern "C" int art_gram_sse(const double* grams, int K, int p, int bias_col, const double* beta,
                            double* sse, double* wsum, hipStream_t stream) {
  if (K < 0 || p < 1 || p > 128 || bias_col >= p || !sse || (K > 0 && (!grams || !beta)))
    return fail(ART_ERR_ARG, "art_gram_sse: bad arguments");
  if (K == 0) return ART_OK;
  const int T = (p + 1 + 15) / 16, npairs = T * (T + 1) / 2;
  gram_sse_kernel<<<K, 256, 0, stream>>>(grams, T, npairs, p, bias_col, beta, sse, wsum);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "gram_sse_kernel launch");
}

extern "C" int art_gather_strided(const double* X, const double* y, const double* w, int d,
                                  int K, const int64_t* src_begin, const int64_t* src_end,
                                  const int64_t* dst_offsets, int64_t total, double* Xo,
                                  double* yo, double* wo, hipStream_t stream) {
  if (d < 1 || K < 0 || total < 0 || (w == nullptr) != (wo == nullptr))
    return fail(ART_ERR_ARG, "art_gather_strided: bad arguments");
  if (K == 0 || total == 0) return ART_OK;
  if (!X || !y || !src_begin || !src_end || !dst_offsets || !Xo || !yo)
    return fail(ART_ERR_ARG, "art_gather_strided: null pointer");
  const unsigned blocks = (unsigned)((total + 255) / 256);
  gather_strided_kernel<<<blocks, 256, 0, stream>>>(X, y, w, d, K, src_begin, src_end,
                                                    dst_offsets, total, Xo, yo, wo);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "gather_strided_kernel launch");
}

extern "C" int art_partition_rows(const double* X, const double* y, const double* w, int d,
                                  int K, const int64_t* src_begin, const int64_t* src_end,
                                  const int64_t* blk_offsets,
                                  int64_t blocks, const int64_t* dst_base, const double* node_w,
                                  const double* node_z, unsigned char* flags, int* blk_right,
                                  int64_t* node_right, double* Xo, double* yo, double* wo,
                                  hipStream_t stream) {
  if (d < 1 || K < 0 || blocks < 0 || (w == nullptr) != (wo == nullptr))
    return fail(ART_ERR_ARG, "art_partition_rows: bad arguments");
  if (K == 0) return ART_OK;
  if (!src_begin || !src_end || !blk_offsets || !dst_base || !node_w || !node_z || !node_right ||
      (blocks > 0 && (!X || !y || !flags || !blk_right || !Xo || !yo)))
    return fail(ART_ERR_ARG, "art_partition_rows: null pointer");
  if (blocks > 0x7fffffff) return fail(ART_ERR_ARG, "art_partition_rows: too many blocks");
  PartArgs a{X, y, w, d, K, src_begin, src_end, blk_offsets, dst_base, node_w, node_z,
             flags, blk_right, node_right, Xo, yo, wo};
  if (blocks > 0) partition_count_kernel<<<(unsigned)blocks, kPartThreads, 0, stream>>>(a);
  partition_scan_kernel<<<(unsigned)((K + 255) / 256), 256, 0, stream>>>(a);
  if (blocks > 0) partition_scatter_kernel<<<(unsigned)blocks, kPartThreads, 0, stream>>>(a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "partition kernels launch");
}

extern "C" int art_partition_block_rows(void) { return kPartRows; }
