// Tree inference and node temperature scales.
//
// art_tree_predict: batched prediction of a fitted regression tree
// (art/tree.py:124-165 on the CPU).  The tree is flattened breadth-first
// into split rows (w, z, two child links) and leaf coefficient rows; a lane
// routes one point from the root (right when w.x - z >= 0) and evaluates its
// leaf's ridge model on the fused feature basis.
//
// art_knn_median: median nearest-neighbour distance (or median pairwise
// distance) of many small point sets, the scale that temperature.py:73-137
// turns into a sigmoid temperature.  One workgroup per set: nearest-neighbour
// distances are staged in LDS (brute force over the set), and the median is a
// radix select over order-preserving 64-bit keys; the pairwise mode recomputes
// its n(n-1)/2 distances on every select pass instead of storing them.

#include "art_common.h"
#include "art_hip.h"

namespace art {

namespace {

constexpr int kKnnThreads = 256;
constexpr int kKnnMaxPoints = 4096;

// Leaf-model value at x (registers x[0..D)) for the tree's basis.
template <int D>
__device__ inline double leaf_value(const double (&x)[D], const Basis& B,
                                    const double* __restrict__ c) {
  const int d = B.d;
  double acc = 0.0;
  int k = 0;
  if (B.bias && !B.bias_last) acc = c[k++];
#pragma unroll
  for (int j = 0; j < D; ++j)
    if (j < d) acc = fma(x[j], c[k++], acc);
  if (B.degree == 2) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      if (i >= d) break;
#pragma unroll
      for (int j = i; j < D; ++j)
        if (j < d) acc = fma(x[i] * x[j], c[k++], acc);
    }
  }
  if (B.bias && B.bias_last) acc += c[k];
  return acc;
}

template <int D>
__global__ __launch_bounds__(256) void tree_predict_kernel(
    const double* __restrict__ X, int64_t n, const double* __restrict__ node_w,
    const double* __restrict__ node_z, const int* __restrict__ node_child, int root,
    const double* __restrict__ leaf_coef, Basis B, double* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int d = B.d;
  double x[D];
#pragma unroll
  for (int j = 0; j < D; ++j) x[j] = j < d ? X[r * d + j] : 0.0;
  int node = root;
  while (node >= 0) {
    const double* w = node_w + (size_t)node * d;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j)
      if (j < d) s = fma(w[j], x[j], s);
    node = node_child[2 * node + ((s - node_z[node]) >= 0.0 ? 1 : 0)];
  }
  out[r] = leaf_value<D>(x, B, leaf_coef + (size_t)(-1 - node) * B.p);
}

template <int D>
__global__ __launch_bounds__(256) void tree_leaf_kernel(
    const double* __restrict__ X, int64_t n, int d, const double* __restrict__ node_w,
    const double* __restrict__ node_z, const int* __restrict__ node_child, int root,
    int* __restrict__ leaf) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double x[D];
#pragma unroll
  for (int j = 0; j < D; ++j) x[j] = j < d ? X[r * d + j] : 0.0;
  int node = root;
  while (node >= 0) {
    const double* w = node_w + (size_t)node * d;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j)
      if (j < d) s = fma(w[j], x[j], s);
    node = node_child[2 * node + ((s - node_z[node]) >= 0.0 ? 1 : 0)];
  }
  leaf[r] = -1 - node;
}

__device__ inline uint64_t dist_key(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ inline double key_dist(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

__device__ inline double point_distance(const double* a, const double* b, int d) {
  double s = 0.0;
  for (int j = 0; j < d; ++j) {
    const double t = a[j] - b[j];
    s = fma(t, t, s);
  }
  return sqrt(s);
}

struct KnnShared {
  double nn[kKnnMaxPoints];
  unsigned hist[256];
  unsigned long long prefix, mask, above;
  unsigned rank, bin_count;
};

// Key source of one set: staged nearest-neighbour distances (mode 0) or the
// pairwise distances of the points (mode 1, recomputed on each call).
struct SetKeys {
  const double* P;
  const double* nn;
  int n, d, mode;

  template <typename F>
  __device__ void each(F&& f) const {
    if (mode == 0) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) f(dist_key(nn[i]));
    } else {
      for (int i = threadIdx.x; i < n; i += blockDim.x)
        for (int j = i + 1; j < n; ++j) f(dist_key(point_distance(P + (size_t)i * d,
                                                                  P + (size_t)j * d, d)));
    }
  }
};

// The rank-th smallest key: S.prefix, its rank among equal keys S.rank and
// the number of equal keys S.bin_count.
__device__ void knn_select(const SetKeys& keys, KnnShared& S, unsigned rank) {
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
    keys.each([&](unsigned long long key) {
      if ((key & mask) == prefix) atomicAdd(&S.hist[(key >> shift) & 255u], 1u);
    });
    __syncthreads();
    if (tid < 64) {
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

__global__ __launch_bounds__(kKnnThreads) void knn_median_kernel(
    const double* __restrict__ X, const int64_t* __restrict__ offsets, int d, int mode,
    double* __restrict__ out) {
  __shared__ KnnShared S;
  const int set = blockIdx.x, tid = threadIdx.x;
  const int64_t begin = offsets[set];
  const int n = (int)(offsets[set + 1] - begin);
  const double* P = X + begin * d;
  if (mode == 0) {
    for (int i = tid; i < n; i += blockDim.x) {
      double best = __builtin_inf();
      const double* a = P + (size_t)i * d;
      for (int j = 0; j < n; ++j)
        if (j != i) best = fmin(best, point_distance(a, P + (size_t)j * d, d));
      S.nn[i] = best;
    }
  }
  const SetKeys keys{P, S.nn, n, d, mode};
  const unsigned long long count =
      mode == 0 ? (unsigned long long)n : (unsigned long long)n * (n - 1) / 2;
  // knn_select starts with a barrier, which also publishes S.nn.
  knn_select(keys, S, (unsigned)((count - 1) / 2));
  const unsigned long long lo = S.prefix;
  double med = key_dist(lo);
  if (!(count & 1)) {
    const bool repeats = S.rank + 1 < S.bin_count;
    unsigned long long hi = lo;
    if (!repeats) {
      if (tid == 0) S.above = ~0ull;
      __syncthreads();
      unsigned long long best = ~0ull;
      keys.each([&](unsigned long long key) {
        if (key > lo && key < best) best = key;
      });
      if (best != ~0ull) atomicMin(&S.above, best);
      __syncthreads();
      hi = S.above;
    }
    med = (med + key_dist(hi)) * 0.5;
  }
  if (tid == 0) out[set] = med;
}

}  // namespace

}  // namespace art

using namespace art;

extern "C" int art_tree_predict(const double* X, int64_t n, const double* node_w,
                                const double* node_z, const int* node_child, int root,
                                const double* leaf_coef, art_basis basis, double* out,
                                hipStream_t stream) {
  if (n < 0 || basis.d < 1 || basis.d > 32 || basis.p < 1 || (basis.degree != 1 && basis.degree != 2))
    return fail(ART_ERR_ARG, "art_tree_predict: needs 1 <= d <= 32 and a degree 1 or 2 basis");
  if (basis.p != basis_size(basis.d, basis.degree, basis.bias))
    return fail(ART_ERR_ARG, "art_tree_predict: basis.p does not match d, degree and bias");
  if (root >= 0 && (!node_w || !node_z || !node_child))
    return fail(ART_ERR_ARG, "art_tree_predict: null split arrays");
  if (n == 0) return ART_OK;
  if (!X || !leaf_coef || !out) return fail(ART_ERR_ARG, "art_tree_predict: null pointer");
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (basis.d <= 4)
    tree_predict_kernel<4><<<blocks, 256, 0, stream>>>(X, n, node_w, node_z, node_child, root,
                                                       leaf_coef, basis, out);
  else if (basis.d <= 8)
    tree_predict_kernel<8><<<blocks, 256, 0, stream>>>(X, n, node_w, node_z, node_child, root,
                                                       leaf_coef, basis, out);
  else if (basis.d <= 16)
    tree_predict_kernel<16><<<blocks, 256, 0, stream>>>(X, n, node_w, node_z, node_child, root,
                                                        leaf_coef, basis, out);
  else
    tree_predict_kernel<32><<<blocks, 256, 0, stream>>>(X, n, node_w, node_z, node_child, root,
                                                        leaf_coef, basis, out);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "tree_predict_kernel launch");
}

extern "C" int art_knn_median(const double* X, const int64_t* offsets, const int64_t* host_offsets,
                              int sets, int d, int mode, double* out, hipStream_t stream) {
  if (sets < 0 || d < 1 || (mode != 0 && mode != 1) || !host_offsets)
    return fail(ART_ERR_ARG, "art_knn_median: bad arguments (mode 0 or 1, d >= 1)");
  if (sets == 0) return ART_OK;
  for (int s = 0; s < sets; ++s) {
    const int64_t n = host_offsets[s + 1] - host_offsets[s];
    if (n < 2 || n > kKnnMaxPoints)
      return fail(ART_ERR_ARG, "art_knn_median: every set needs 2 to 4096 points");
  }
  if (!X || !offsets || !out) return fail(ART_ERR_ARG, "art_knn_median: null pointer");
  knn_median_kernel<<<sets, kKnnThreads, 0, stream>>>(X, offsets, d, mode, out);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "knn_median_kernel launch");
}

extern "C" int art_tree_leaf(const double* X, int64_t n, int d, const double* node_w,
                             const double* node_z, const int* node_child, int root, int* leaf,
                             hipStream_t stream) {
  if (n < 0 || d < 1 || d > 32) return fail(ART_ERR_ARG, "art_tree_leaf: needs 1 <= d <= 32");
  if (root >= 0 && (!node_w || !node_z || !node_child))
    return fail(ART_ERR_ARG, "art_tree_leaf: null split arrays");
  if (n == 0) return ART_OK;
  if (!X || !leaf) return fail(ART_ERR_ARG, "art_tree_leaf: null pointer");
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (d <= 8)
    tree_leaf_kernel<8><<<blocks, 256, 0, stream>>>(X, n, d, node_w, node_z, node_child, root, leaf);
  else if (d <= 16)
    tree_leaf_kernel<16><<<blocks, 256, 0, stream>>>(X, n, d, node_w, node_z, node_child, root,
                                                     leaf);
  else
    tree_leaf_kernel<32><<<blocks, 256, 0, stream>>>(X, n, d, node_w, node_z, node_child, root,
                                                     leaf);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "tree_leaf_kernel launch");
}
