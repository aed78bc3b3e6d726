// Shared device helpers for the ART hot-path kernels (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace art {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kWave = 64;

// Feature-basis descriptor shared by every kernel that builds leaf features
// from raw inputs.  Column order (art/_cuda_features.py:13-29):
//   degree 2: [1?, x_0..x_{d-1}, x_i*x_j for i<=j with i outermost]
//   degree 1: [1?, x_0..x_{d-1}] or, with bias_last, [x_0..x_{d-1}, 1]
struct Basis {
  int d;
  int p;
  int degree;
  int bias;       // include_bias
  int bias_last;  // affine order (degree 1 only)
};

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

// Row range of one Gram work item (a chunk of one leaf's rows).
struct LeafChunk {
  int64_t begin;
  int64_t end;
  int leaf;
  int pad;
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

}  // namespace art
