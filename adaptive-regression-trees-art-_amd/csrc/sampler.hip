// Hit-and-run sampling of many convex polytopes {x : A x <= b} (row a7).
//
// One lane per chain, one 64-lane workgroup per (polytope, block of 64
// chains).  Per move a lane draws a direction, recomputes every slack
// s_i = b_i - A_i.x exactly in FP64 together with A_i.d (the constraint rows
// are uniform across the wave and arrive through scalar loads), takes the
// chord [max_{A_i.d<0} s_i/A_i.d, min_{A_i.d>0} s_i/A_i.d] and moves to a
// uniform point of it (art/sampling.py:877-962, art/_numba_sampling.py:9-40).
// Recomputing the slack instead of updating it incrementally needs no
// per-chain slack storage and keeps rounding from accumulating along a chain.
// Ratios use a reciprocal (v_rcp_f64 and two Newton steps) times the slack, a
// few ulp from the quotient.  Feasibility by construction: the point is drawn
// from the chord shrunk by kChordMargin of its length at each end,
// t = (t_lo + m) + u ((t_hi - m) - (t_lo + m)), m = kChordMargin (t_hi - t_lo),
// so a binding row keeps a slack of ~2^-40 of the chord times |A_i.d|, far
// above the few-ulp error of the ratios and of evaluating A x - b; the
// sampled distribution moves by 2^-40 of the chord length.
//
// Schedule (one rule for the reference's batched and scalar samplers): chain
// c records after moves first, first + stride, ..., K_c records in all, with
// K_c = n / C + (c < n % C); record k of chain c is sample k*C + c
// (art/sampling.py:745-803 batched: first = burn_in + stride;
// art/sampling.py:480-494 scalar: one chain, first = burn_in + 1).
//
// Random stream (no reference counterpart; the reference draws host NumPy
// arrays): Philox4x32-10 (Salmon et al., SC'11) with key (seed_lo, seed_hi)
// and counter (move * W + w, chain, polytope, 0); FP32 Box-Muller directions
// and a 53-bit FP64 chord uniform.  A fixed-stream mode reads directions and
// uniforms from device arrays instead, for parity with the reference's
// NumPy-driven samplers.

#include "art_common.h"
#include "art_hip.h"

namespace art {
// Relative margin kept inside each end of a chord (2^-40).
constexpr double kChordMargin = 0x1p-40;

// The point of the chord [tlo, thi] shrunk by kChordMargin of its length at
// each end: tlo + (thi - tlo) u' with u' = m + u (1 - 2 m) (chord_fraction,
// taken once per move, off the chain's dependent path), which is
// lo + u (hi - lo) for lo = tlo + m (thi - tlo), hi = thi - m (thi - tlo).
// 1 - 2 m is exact, and u' lies in [m, 1 - m] for u in [0, 1).
__device__ __forceinline__ double chord_fraction(double u) {
  return fma(u, 1.0 - 2.0 * kChordMargin, kChordMargin);
}
__device__ __forceinline__ double chord_point(double tlo, double thi, double uf) {
  return fma(thi - tlo, uf, tlo);
}
}  // namespace art

#include <utility>

namespace art {

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // Each 64-bit product is one v_mad_u64_u32 (separate low and high 32-bit
    // multiplies would be two quarter-rate instructions).
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}

__host__ __device__ constexpr int har_calls_per_step(int d) {
  return (2 * ((d + 1) / 2) + 2 + 3) / 4;
}

__global__ __launch_bounds__(256) void philox_kernel(const uint4* __restrict__ ctr,
                                                     const uint2* __restrict__ key, int64_t count,
                                                     uint4* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = philox4x32_10(ctr[i], key[i].x, key[i].y);
}

// Read-only inputs go through the constant address space: uniform indices
// then compile to scalar loads (one copy per wave, no vector-memory waits).
typedef const __attribute__((address_space(4))) double cf64;
typedef const __attribute__((address_space(4))) int64_t ci64;

// Direction (shaped) and chord uniform of move s of chain c: the Philox
// stream or the fixed arrays.  Both kernels call this, so a chain draws the
// same bits whichever kernel runs it.
template <int DM, bool EXACT>
__device__ __forceinline__ void move_randoms(const art_har_args& a, cf64* shape, int d, int W,
                                             int poly, uint32_t stream_poly, int c, int s,
                                             uint32_t k0, uint32_t k1, double (&dir)[DM],
                                             double& u) {
  constexpr int WM = har_calls_per_step(DM);
  const int C = a.chains;
  if (a.directions != nullptr) {
    const size_t row = ((size_t)poly * a.steps + s) * C + c;
    const double* dp = a.directions + row * d;
#pragma unroll
    for (int j = 0; j < DM; ++j) dir[j] = (EXACT || j < d) ? dp[j] : 0.0;
    u = a.uniforms[row];
  } else {
    uint32_t w[4 * WM];
#pragma unroll
    for (int q = 0; q < WM; ++q) {
      uint4 r = make_uint4(0u, 0u, 0u, 0u);
      if (EXACT || q < W)
        r = philox4x32_10(make_uint4((uint32_t)(s * W + q), (uint32_t)c, stream_poly, 0u), k0,
                          k1);
      w[4 * q] = r.x;
      w[4 * q + 1] = r.y;
      w[4 * q + 2] = r.z;
      w[4 * q + 3] = r.w;
    }
    // Box-Muller pairs from words (2q, 2q+1): u1 in (0, 1] keeps the log
    // finite; the hardware sine and cosine take revolutions, so
    // sin(2 pi u2) is v_sin_f32(u2).
#pragma unroll
    for (int q = 0; q < (DM + 1) / 2; ++q) {
      const float u1 = (float)((w[2 * q] >> 8) + 1u) * 0x1p-24f;
      const float u2 = (float)(w[2 * q + 1] >> 8) * 0x1p-24f;
      const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
      dir[2 * q] = (double)(r * __builtin_amdgcn_cosf(u2));
      if (2 * q + 1 < DM) dir[2 * q + 1] = (double)(r * __builtin_amdgcn_sinf(u2));
    }
    // The chord uniform follows the ceil(d/2) pairs.
    uint32_t wa, wb;
    if (EXACT) {
      wa = w[2 * ((DM + 1) / 2)];
      wb = w[2 * ((DM + 1) / 2) + 1];
    } else {
      const int P = (d + 1) / 2;
      wa = wb = 0u;
#pragma unroll
      for (int q = 0; q < (DM + 1) / 2; ++q)
        if (q == P - 1) {
          wa = w[2 * q + 2];
          wb = w[2 * q + 3];
        }
    }
    u = ((double)(wa >> 5) * 67108864.0 + (double)(wb >> 6)) * 0x1p-53;
    if (!EXACT) {
#pragma unroll
      for (int j = 0; j < DM; ++j)
        if (j >= d) dir[j] = 0.0;
    }
  }
  if (shape) {
    double t[DM];
#pragma unroll
    for (int i = 0; i < DM; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < DM; ++j)
        if (EXACT || (i < d && j < d)) acc = fma(shape[i * d + j], dir[j], acc);
      t[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < DM; ++i) dir[i] = t[i];
  }
}

__device__ __forceinline__ double rcp2(double v) {
  double r = __builtin_amdgcn_rcp(v);
  r = fma(r, fma(-v, r, 1.0), r);
  r = fma(r, fma(-v, r, 1.0), r);
  return r;
}

// Chain c of polytope poly, one lane: DM = the compile-time dimension;
// EXACT: d == DM (no per-column predicates).
template <int DM, bool EXACT>
__device__ __attribute__((always_inline)) inline void har_chain_lane(const art_har_args& a, int poly, int c) {
  const int d = EXACT ? DM : a.d;
  const int C = a.chains;
  ci64* offs = (ci64*)a.row_offsets;
  const int64_t r0 = offs[poly];
  const int m = (int)(offs[poly + 1] - r0);
  cf64* A = (cf64*)(a.A + r0 * d);
  cf64* b = (cf64*)(a.b + r0);
  cf64* shape = (cf64*)(a.shape ? a.shape + (size_t)poly * a.shape_stride : nullptr);
  cf64* box_lo = (cf64*)(a.box_lo ? a.box_lo + (size_t)poly * d : nullptr);
  cf64* box_hi = (cf64*)(a.box_hi ? a.box_hi + (size_t)poly * d : nullptr);
  const int full = a.n / C, rem = a.n % C;
  const int records = full + (c < rem ? 1 : 0);
  const int moves = a.first + (records - 1) * a.stride;
  const int W = har_calls_per_step(d);
  const uint64_t key = a.seeds ? a.seeds[poly] : a.seed;
  const uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
  const uint32_t stream_poly = a.seeds ? 0u : (uint32_t)poly;

  double x[DM];
  const double* x0 = a.x0 + ((size_t)poly * C + c) * d;
#pragma unroll
  for (int j = 0; j < DM; ++j) x[j] = (EXACT || j < d) ? x0[j] : 0.0;
  double* out = a.samples + (size_t)poly * a.n * d;

  for (int s = 0; s < moves; ++s) {
    double dir[DM];
    double u;
    move_randoms<DM, EXACT>(a, shape, d, W, poly, stream_poly, c, s, k0, k1, dir, u);
    double tlo = -INFINITY, thi = INFINITY;
    if (box_lo) {
      // Axis rows x_j <= hi_j and -x_j <= -lo_j: slack and direction product
      // need no multiplication, and the tighter of repeated rows always wins
      // the chord's min/max, so one bound per side gives the same chord.  The
      // two ratios share one reciprocal (v_rcp_f64 plus two Newton steps, a
      // few ulp from the quotient), and the direction's sign says which face
      // bounds which end.  A zero component (reciprocal +-inf, the same sign)
      // gives the ends -inf and +inf, or NaN for a point on a face, which
      // fmax / fmin drop: no bound, as for the skipped component of the
      // single-chain replay.
#pragma unroll
      for (int j = 0; j < DM; ++j)
        if (EXACT || j < d) {
          const double dj = dir[j];
          double r = __builtin_amdgcn_rcp(dj);
          r = fma(r, fma(-dj, r, 1.0), r);
          r = fma(r, fma(-dj, r, 1.0), r);
          const double rl = (box_lo[j] - x[j]) * r;
          const double rh = (box_hi[j] - x[j]) * r;
          const bool up = !__builtin_signbit(dj);
          tlo = fmax(tlo, up ? rl : rh);
          thi = fmin(thi, up ? rh : rl);
        }
    }
#pragma unroll 2
    for (int i = 0; i < m; ++i) {
      double ad = 0.0, ax = 0.0;
#pragma unroll
      for (int j = 0; j < DM; ++j)
        if (EXACT || j < d) {
          const double aij = A[i * d + j];
          ad = fma(aij, dir[j], ad);
          ax = fma(aij, x[j], ax);
        }
      double r = __builtin_amdgcn_rcp(ad);
      r = fma(r, fma(-ad, r, 1.0), r);
      r = fma(r, fma(-ad, r, 1.0), r);
      const double ratio = (b[i] - ax) * r;
      tlo = (ad < 0.0 && ratio > tlo) ? ratio : tlo;
      thi = (ad > 0.0 && ratio < thi) ? ratio : thi;
    }
    const double t = chord_point(tlo, thi, chord_fraction(u));
#pragma unroll
    for (int j = 0; j < DM; ++j) x[j] += dir[j] * t;
    const int done = s + 1;
    if (done >= a.first && (done - a.first) % a.stride == 0) {
      double* o = out + ((size_t)((done - a.first) / a.stride) * C + c) * d;
#pragma unroll
      for (int j = 0; j < DM; ++j)
        if (EXACT || j < d) o[j] = x[j];
    }
  }
}

// A lane per chain, a wave per 64 chains of one polytope.
template <int DM, bool EXACT>
__global__ __launch_bounds__(64) void har_kernel(art_har_args a) {
  const int wpp = (a.chains + 63) / 64;
  const int poly = blockIdx.x / wpp;
  const int c = (blockIdx.x - poly * wpp) * 64 + (int)threadIdx.x;
  if (c < a.chains) har_chain_lane<DM, EXACT>(a, poly, c);
}

// One chain per polytope (the scalar sampler): the wave splits each block
// of 64 moves.  Lane l first draws move s0 + l's direction and uniform and
// the move's state-free quantities -- A_i.d with its reciprocal for every
// row and the reciprocal of each direction component -- into LDS; then
// every lane replays the 64 moves of the one chain from LDS (the same values
// in every lane), so a move's critical path is only the slacks, the chord
// and the step.  Each quantity is computed by the same code as in
// har_kernel, so the chain is bit-identical to har_kernel's.
// LDS entry of a move: dir[DM], 1/dir[DM], the chord fraction u', A_i.d[m], 1/(A_i.d)[m].
constexpr int kChainBlock = 64;

// One chain per workgroup, M >= 0: exactly M general rows, held in registers
// with b for the whole chain (the replay then reads no constraint data per
// move; a scalar load on every row of every move otherwise); M < 0: any row
// count, read per move.  The rows are scanned in the same order either way,
// so both forms give the same chain.
template <int DM, bool EXACT, int M>
__device__ __attribute__((always_inline)) inline void chain_run(const art_har_args& a, int E,
                                                               double* chain_lds) {
  const int d = EXACT ? DM : a.d;
  const int poly = blockIdx.x, lane = (int)threadIdx.x;
  ci64* offs = (ci64*)a.row_offsets;
  const int64_t r0 = offs[poly];
  const int m = (int)(offs[poly + 1] - r0);
  cf64* A = (cf64*)(a.A + r0 * d);
  cf64* b = (cf64*)(a.b + r0);
  cf64* shape = (cf64*)(a.shape ? a.shape + (size_t)poly * a.shape_stride : nullptr);
  cf64* box_lo = (cf64*)(a.box_lo ? a.box_lo + (size_t)poly * d : nullptr);
  cf64* box_hi = (cf64*)(a.box_hi ? a.box_hi + (size_t)poly * d : nullptr);
  const int moves = a.first + (a.n - 1) * a.stride;
  const int W = har_calls_per_step(d);
  const uint64_t key = a.seeds ? a.seeds[poly] : a.seed;
  const uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
  const uint32_t stream_poly = a.seeds ? 0u : (uint32_t)poly;

  double x[DM];
  const double* x0 = a.x0 + (size_t)poly * d;
#pragma unroll
  for (int j = 0; j < DM; ++j) x[j] = (EXACT || j < d) ? x0[j] : 0.0;
  double* out = a.samples + (size_t)poly * a.n * d;
  // Box bounds in registers for the whole chain, and a countdown to the next
  // recorded move instead of a division per move.
  double blo[DM], bhi[DM];
#pragma unroll
  for (int j = 0; j < DM; ++j) {
    blo[j] = (box_lo && (EXACT || j < d)) ? box_lo[j] : 0.0;
    bhi[j] = (box_lo && (EXACT || j < d)) ? box_hi[j] : 0.0;
  }
  int until_record = a.first, recorded = 0;
  constexpr int MR = M > 0 ? M : 1;
  constexpr int MR0 = M > 0 ? M : 0;
  double Ar[MR][DM], br[MR];
  if constexpr (M > 0) {
#pragma unroll
    for (int i = 0; i < M; ++i) {
      br[i] = b[i];
#pragma unroll
      for (int j = 0; j < DM; ++j) Ar[i][j] = (EXACT || j < d) ? A[i * d + j] : 0.0;
    }
  }

  // Doubles per move entry: the launch's bound E for the per-move row
  // reads; with M rows in registers, this polytope's own 4 DM + 1 + 2 M, a
  // compile-time stride (the workgroup's entries are its own).
  const int ES = M >= 0 ? 4 * DM + 1 + 2 * MR0 : E;
  for (int s0 = 0; s0 < moves; s0 += kChainBlock) {
    const int cnt = min(kChainBlock, moves - s0);
    if (lane < cnt) {
      double dir[DM], u;
      move_randoms<DM, EXACT>(a, shape, d, W, poly, stream_poly, 0, s0 + lane, k0, k1, dir, u);
      double* e = chain_lds + (size_t)lane * ES;
      // Per coordinate: the box face the chord leaves through on each side
      // is fixed by the direction's sign, so the move stores it (or +-inf
      // with a unit scale when the coordinate does not move or there is no
      // box: the replay's (face - x) * scale is then -inf / +inf).
#pragma unroll
      for (int j = 0; j < DM; ++j)
        if (EXACT || j < d) {
          const bool nz = box_lo && dir[j] != 0.0, up = dir[j] > 0.0;
          e[j] = dir[j];
          e[DM + j] = nz ? rcp2(dir[j]) : 1.0;
          e[2 * DM + 1 + j] = nz ? (up ? blo[j] : bhi[j]) : -INFINITY;
          e[3 * DM + 1 + j] = nz ? (up ? bhi[j] : blo[j]) : INFINITY;
        }
      e[2 * DM] = chord_fraction(u);
      // Per row, 1 / (A_i.d) on the side of the chord it bounds and NaN on
      // the other (a row with A_i.d = 0 bounds neither): the replay's
      // slack * reciprocal is then the row's chord end or NaN, and fmax /
      // fmin (IEEE maxNum / minNum) drop the NaNs.
      double* ea = e + 4 * DM + 1;
      for (int i = 0; i < m; ++i) {
        double ad = 0.0;
#pragma unroll
        for (int j = 0; j < DM; ++j)
          if (EXACT || j < d) ad = fma(A[i * d + j], dir[j], ad);
        const double r = rcp2(ad);
        ea[i] = ad < 0.0 ? r : NAN;
        ea[m + i] = ad > 0.0 ? r : NAN;
      }
    }
    __syncthreads();
    if constexpr (M >= 0) {
      // Rows in registers: the replay of move k works from registers only.
      // Move k + 1's entry is read from LDS while move k runs (its reads do
      // not depend on the chain), and the chord ends are fmax / fmin trees
      // over the candidates (exact, so the values of a running max / min,
      // without a compare and select per candidate).
      constexpr int NC = DM + MR;
      struct Entry {
        double dir[DM], r[DM], flo[DM], fhi[DM], u, rlo[MR], rhi[MR];
      };
      auto read = [&](int k, Entry& q) {
        const double* e = chain_lds + (size_t)k * ES;
#pragma unroll
        for (int j = 0; j < DM; ++j) {
          const bool on = EXACT || j < d;
          q.dir[j] = on ? e[j] : 0.0;
          q.r[j] = on ? e[DM + j] : 1.0;
          q.flo[j] = on ? e[2 * DM + 1 + j] : -INFINITY;
          q.fhi[j] = on ? e[3 * DM + 1 + j] : INFINITY;
        }
        q.u = e[2 * DM];
#pragma unroll
        for (int i = 0; i < M; ++i) {
          q.rlo[i] = e[4 * DM + 1 + i];
          q.rhi[i] = e[4 * DM + 1 + M + i];
        }
      };
      auto step = [&](const Entry& cur) __attribute__((always_inline)) {
        double lo[NC], hi[NC];
#pragma unroll
        for (int j = 0; j < DM; ++j) {
          if (EXACT || j < d) {
            lo[j] = (cur.flo[j] - x[j]) * cur.r[j];
            hi[j] = (cur.fhi[j] - x[j]) * cur.r[j];
          } else {
            lo[j] = -INFINITY;
            hi[j] = INFINITY;
          }
        }
#pragma unroll
        for (int i = 0; i < MR; ++i) {
          if (i < M) {
            double ax = 0.0;
#pragma unroll
            for (int j = 0; j < DM; ++j)
              if (EXACT || j < d) ax = fma(Ar[i][j], x[j], ax);
            const double slack = br[i] - ax;
            lo[DM + i] = slack * cur.rlo[i];
            hi[DM + i] = slack * cur.rhi[i];
          } else {
            lo[DM + i] = -INFINITY;
            hi[DM + i] = INFINITY;
          }
        }
#pragma unroll
        for (int w = 1; w < NC; w *= 2)
#pragma unroll
          for (int i = 0; i + w < NC; i += 2 * w) {
            lo[i] = fmax(lo[i], lo[i + w]);
            hi[i] = fmin(hi[i], hi[i + w]);
          }
        const double t = chord_point(lo[0], hi[0], cur.u);
#pragma unroll
        for (int j = 0; j < DM; ++j) x[j] += cur.dir[j] * t;
        if (--until_record == 0) {  // moves first, first + stride, ...
          if (lane == 0) {
            double* o = out + (size_t)recorded * d;
#pragma unroll
            for (int j = 0; j < DM; ++j)
              if (EXACT || j < d) o[j] = x[j];
          }
          ++recorded;
          until_record = a.stride;
        }
      };
      // Two entries in turn (no copy of a loop-carried entry per move): the
      // next move's entry loads while the current move runs.
      Entry e0, e1;
      read(0, e0);
      int k = 0;
      for (; k + 2 <= cnt; k += 2) {
        read(k + 1, e1);
        step(e0);
        read(k + 2 < cnt ? k + 2 : k + 1, e0);
        step(e1);
      }
      if (k < cnt) step(e0);
      __syncthreads();
      continue;
    }
    for (int k = 0; k < cnt; ++k) {
      const double* e = chain_lds + (size_t)k * E;
      double tlo = -INFINITY, thi = INFINITY;
      if (box_lo) {
        // (face - x) * (1 / d_j): the same products as ordering the two face
        // ratios by value, the order taken from the direction's sign.
#pragma unroll
        for (int j = 0; j < DM; ++j)
          if (EXACT || j < d) {
            const double r = e[DM + j];
            const double lo_c = (e[2 * DM + 1 + j] - x[j]) * r;
            const double hi_c = (e[3 * DM + 1 + j] - x[j]) * r;
            tlo = lo_c > tlo ? lo_c : tlo;
            thi = hi_c < thi ? hi_c : thi;
          }
      }
      const double* ea = e + 4 * DM + 1;
      {  // any row count (M < 0): rows read per move
#pragma unroll 2
        for (int i = 0; i < m; ++i) {
          double ax = 0.0;
#pragma unroll
          for (int j = 0; j < DM; ++j)
            if (EXACT || j < d) ax = fma(A[i * d + j], x[j], ax);
          const double slack = b[i] - ax;
          tlo = fmax(tlo, slack * ea[i]);
          thi = fmin(thi, slack * ea[m + i]);
        }
      }
      const double t = chord_point(tlo, thi, e[2 * DM]);
#pragma unroll
      for (int j = 0; j < DM; ++j) x[j] += e[j] * t;
      if (--until_record == 0) {  // moves first, first + stride, ...
        if (lane == 0) {
          double* o = out + (size_t)recorded * d;
#pragma unroll
          for (int j = 0; j < DM; ++j)
            if (EXACT || j < d) o[j] = x[j];
        }
        ++recorded;
        until_record = a.stride;
      }
    }
    __syncthreads();
  }
}


// Single-chain kernel: polytopes of at most 8 general rows in d <= 4 (the
// builder's nodes: a box and the splits above them) run with their rows in
// registers.
template <int DM, bool EXACT>
__global__ __launch_bounds__(64) void har_chain_kernel(art_har_args a, int E) {
  extern __shared__ double chain_lds[];
  ci64* offs = (ci64*)a.row_offsets;
  const int m = (int)(offs[blockIdx.x + 1] - offs[blockIdx.x]);
  // A move's LDS entry holds 2 (E - 4 DM - 1) / 2 row values: a polytope
  // with more rows than the launch's row bound replays its chain on one lane
  // by the per-chain code (the same chain), instead of overrunning LDS.
  if (m > (E - 4 * DM - 1) / 2) {
    if (threadIdx.x == 0) har_chain_lane<DM, EXACT>(a, (int)blockIdx.x, 0);
    return;
  }
  if constexpr (DM <= 4) {
    switch (m) {
#define ART_ROWS(MM) \
  case MM:           \
    chain_run<DM, EXACT, MM>(a, E, chain_lds); \
    return;
      ART_ROWS(0) ART_ROWS(1) ART_ROWS(2) ART_ROWS(3) ART_ROWS(4) ART_ROWS(5) ART_ROWS(6)
      ART_ROWS(7) ART_ROWS(8)
#undef ART_ROWS
      default:
        break;
    }
  }
  chain_run<DM, EXACT, -1>(a, E, chain_lds);
}

// LDS doubles per move of har_chain_kernel for dimension-bucket DM and m rows.
inline int chain_entry(int DM, int m) { return 4 * DM + 1 + 2 * m; }
constexpr int kChainLdsBytes = 64 * 1024;

template <int DM, bool EXACT>
void launch_chain(const art_har_args& a, int max_rows, hipStream_t stream) {
  const int E = chain_entry(DM, max_rows);
  // At most kChainLdsBytes (64 KB): within the default dynamic LDS limit.
  const size_t bytes = sizeof(double) * (size_t)kChainBlock * E;
  har_chain_kernel<DM, EXACT><<<a.P, 64, bytes, stream>>>(a, E);
}

template <int DM, bool EXACT>
void launch_har(const art_har_args& a, hipStream_t stream) {
  const int blocks = a.P * ((a.chains + 63) / 64);
  har_kernel<DM, EXACT><<<blocks, 64, 0, stream>>>(a);
}

using HarLaunch = void (*)(const art_har_args&, hipStream_t);

template <int... Ds>
constexpr HarLaunch exact_launch(int d, std::integer_sequence<int, Ds...>) {
  HarLaunch table[] = {launch_har<Ds + 1, true>...};
  return table[d - 1];
}

using ChainLaunch = void (*)(const art_har_args&, int, hipStream_t);

template <int... Ds>
constexpr ChainLaunch exact_chain(int d, std::integer_sequence<int, Ds...>) {
  ChainLaunch table[] = {launch_chain<Ds + 1, true>...};
  return table[d - 1];
}

}  // namespace art

using namespace art;

extern "C" int art_philox4x32_10(const uint32_t* counters, const uint32_t* keys, int64_t count,
                                 uint32_t* out, hipStream_t stream) {
  if (count < 0 || (count > 0 && (!counters || !keys || !out)))
    return fail(ART_ERR_ARG, "art_philox4x32_10: bad arguments");
  if (count == 0) return ART_OK;
  const int64_t blocks = (count + 255) / 256;
  philox_kernel<<<(unsigned)blocks, 256, 0, stream>>>((const uint4*)counters, (const uint2*)keys,
                                                      count, (uint4*)out);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "philox_kernel launch");
}

extern "C" int art_har_sample(const art_har_args* a, hipStream_t stream) {
  if (!a) return fail(ART_ERR_ARG, "art_har_sample: null arguments");
  if (a->P < 1 || a->d < 1 || a->d > 32 || a->chains < 1 || a->n < a->chains || a->first < 1 ||
      a->stride < 1)
    return fail(ART_ERR_ARG,
                "art_har_sample: need P >= 1, 1 <= d <= 32, 1 <= chains <= n, first >= 1, "
                "stride >= 1");
  if (!a->row_offsets || !a->x0 || !a->samples)
    return fail(ART_ERR_ARG, "art_har_sample: null device pointer");
  if ((a->box_lo == nullptr) != (a->box_hi == nullptr))
    return fail(ART_ERR_ARG, "art_har_sample: box_lo and box_hi go together");
  if ((a->directions == nullptr) != (a->uniforms == nullptr))
    return fail(ART_ERR_ARG, "art_har_sample: directions and uniforms go together");
  if (a->shape_stride < 0 || (a->shape_stride > 0 && (!a->shape || a->shape_stride < a->d * a->d)))
    return fail(ART_ERR_ARG, "art_har_sample: shape_stride must be 0 or at least d*d with a shape");
  const int full = a->n / a->chains, rem = a->n % a->chains;
  const int64_t moves = a->first + (int64_t)(full + (rem ? 1 : 0) - 1) * a->stride;
  if (moves > (int64_t)0x7fffffff / har_calls_per_step(32))
    return fail(ART_ERR_ARG, "art_har_sample: too many moves per chain");
  if (a->directions && a->steps < moves)
    return fail(ART_ERR_ARG, "art_har_sample: fixed stream shorter than the longest chain");
  if ((int64_t)a->P * ((a->chains + 63) / 64) > 0x7fffffff)
    return fail(ART_ERR_ARG, "art_har_sample: grid too large");
  const int d = a->d;
  const int bucket = d <= 16 ? d : (d <= 24 ? 24 : 32);
  const int rows = a->row_bound - 1;
  if (a->chains == 1 && rows >= 0 &&
      sizeof(double) * kChainBlock * chain_entry(bucket, rows) <= (size_t)kChainLdsBytes) {
    if (d <= 16) exact_chain(d, std::make_integer_sequence<int, 16>{})(*a, rows, stream);
    else if (d <= 24) launch_chain<24, false>(*a, rows, stream);
    else launch_chain<32, false>(*a, rows, stream);
  } else if (d <= 16) {
    exact_launch(d, std::make_integer_sequence<int, 16>{})(*a, stream);
  } else if (d <= 24) {
    launch_har<24, false>(*a, stream);
  } else {
    launch_har<32, false>(*a, stream);
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? ART_OK : hip_fail(e, "har_kernel launch");
}
