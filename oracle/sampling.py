"""FP64 oracle for hit-and-run polytope sampling (hot-path row a7).

Test infrastructure only (see ``oracle/__init__.py``).

Restates, with injected random arrays:
* the batched transition `_advance_batch` -- chord [max_{Ad<0} s/Ad,
  min_{Ad>0} s/Ad] with s = b - A x, step t_lo + u (t_hi - t_lo), point and
  incremental-slack update (art/sampling.py:877-962, art/_numba_sampling.py:9-40);
* the batched sampling schedule: burn-in, records after every `stride`
  moves in order round * C + chain, a tail round for the first `remainder`
  chains (art/sampling.py:618-803, art/_numba_sampling.py:43-86);
* the scalar sampler's schedule: a record after moves burn_in + 1,
  burn_in + 1 + thinning, ... (art/sampling.py:445-494);
* direction shaping d <- V diag(sqrt(lambda)) d (art/sampling.py:496-537,865-875);
* the step interval of the scalar sampler (art/sampling.py:539-564).

Plus the device sampler's random stream, which has no reference counterpart
(the reference draws NumPy PCG64/SFC64 arrays on the host):
* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
  easy as 1, 2, 3", SC'11), pinned by the published known-answer vectors;
* word layout: counter (step * W + w, chain, polytope, 0), key (seed low,
  seed high), W = ceil((2 ceil(d/2) + 2) / 4) calls per step;
* directions: FP32 Box-Muller, pair q from words (2q, 2q+1):
  u1 = ((w >> 8) + 1) 2^-24 in (0, 1], u2 = (w >> 8) 2^-24,
  (r cos 2 pi u2, r sin 2 pi u2) with r = sqrt(-2 ln u1);
* chord uniform: words (2P, 2P+1), P = ceil(d/2):
  ((w0 >> 5) 2^26 + (w1 >> 6)) 2^-53 in [0, 1).

Pinned by (tests/test_oracle_sampling.py): the Philox known-answer vectors,
the step-interval known answer of tests/test_sampling.py:446-458, the
scalar == batched fixed-stream equality the reference asserts
(tests/test_sampling.py:175-228), constraint satisfaction of every sample,
and the closed-form moments of the uniform distribution on a box (mean 0,
variance 1/3 on [-1, 1]) and brute-force rejection moments on a cut box.
"""

from __future__ import annotations

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(counter, key):
    """Philox4x32 with 10 rounds; counter (N, 4), key (N, 2) uint32 -> (N, 4)."""
    c = np.array(counter, dtype=np.uint32).reshape(-1, 4).copy()
    k = np.array(key, dtype=np.uint32).reshape(-1, 2).copy()
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k[:, 0] += _W0
                k[:, 1] += _W1
            p0 = _M0 * c[:, 0].astype(np.uint64)
            p1 = _M1 * c[:, 2].astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & _MASK).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & _MASK).astype(np.uint32)
            c = np.stack([hi1 ^ c[:, 1] ^ k[:, 0], lo1, hi0 ^ c[:, 3] ^ k[:, 1], lo0], axis=1)
    return c


def calls_per_step(d):
    return (2 * ((d + 1) // 2) + 2 + 3) // 4


def stream_words(seed, polytope, chain, step, d):
    """The u32 words of one (polytope, chain, step): (..., 4 W) for broadcastable ids."""
    W = calls_per_step(d)
    polytope, chain, step = np.broadcast_arrays(np.asarray(polytope), np.asarray(chain),
                                                np.asarray(step))
    shape = polytope.shape
    w = np.arange(W, dtype=np.uint64)
    ctr = np.zeros(shape + (W, 4), dtype=np.uint32)
    ctr[..., 0] = (step[..., None].astype(np.uint64) * np.uint64(W) + w).astype(np.uint32)
    ctr[..., 1] = chain[..., None].astype(np.uint32)
    ctr[..., 2] = polytope[..., None].astype(np.uint32)
    key = np.empty(shape + (W, 2), dtype=np.uint32)
    key[..., 0] = np.uint32(seed & 0xFFFFFFFF)
    key[..., 1] = np.uint32((seed >> 32) & 0xFFFFFFFF)
    out = philox4x32_10(ctr.reshape(-1, 4), key.reshape(-1, 2))
    return out.reshape(shape + (4 * W,))


def step_randoms(seed, polytope, chain, step, d):
    """FP32 Box-Muller direction (..., d) and FP64 chord uniform (...,)."""
    words = stream_words(seed, polytope, chain, step, d)
    P = (d + 1) // 2
    a = words[..., 0:2 * P:2]
    b = words[..., 1:2 * P:2]
    u1 = ((a >> 8).astype(np.float32) + np.float32(1)) * np.float32(2.0 ** -24)
    u2 = (b >> 8).astype(np.float32) * np.float32(2.0 ** -24)
    r = np.sqrt(np.float32(-2) * np.log(u1))
    ang = 2.0 * np.pi * u2.astype(np.float64)
    cs = np.cos(ang).astype(np.float32)
    sn = np.sin(ang).astype(np.float32)
    direction = np.empty(words.shape[:-1] + (2 * P,), dtype=np.float32)
    direction[..., 0::2] = r * cs
    direction[..., 1::2] = r * sn
    w0 = words[..., 2 * P].astype(np.uint64) >> np.uint64(5)
    w1 = words[..., 2 * P + 1].astype(np.uint64) >> np.uint64(6)
    u = (w0 * np.uint64(1 << 26) + w1).astype(np.float64) * 2.0 ** -53
    return direction[..., :d], u


def step_interval(A, b, x, direction, tol=1e-10):
    """art/sampling.py:539-564 (scalar sampler)."""
    Ax, Ad = A @ x, A @ direction
    slack = b - Ax
    pos, neg = Ad > 1e-14, Ad < -1e-14
    if np.any(~(pos | neg) & (slack < -tol)):
        raise RuntimeError("Current point is infeasible.")
    with np.errstate(divide="ignore", invalid="ignore"):
        r = slack / Ad
    lo = np.max(r[neg], initial=-np.inf)
    hi = np.min(r[pos], initial=np.inf)
    if lo > hi:
        raise RuntimeError("Empty hit-and-run step interval.")
    return float(lo), float(hi)


# The device sampler's chord margin: the point is drawn from the chord shrunk
# by this fraction of its length at each end (feasibility by construction).
# The reference's rule is margin 0.
DEVICE_CHORD_MARGIN = 2.0 ** -40


def chord_point(t_lo, t_hi, u, margin=0.0):
    """The uniform point of the chord shrunk by margin (t_hi - t_lo) at each
    end: t_lo + (t_hi - t_lo) u' with u' = margin + u (1 - 2 margin), i.e.
    lo + u (hi - lo) for lo = t_lo + margin (t_hi - t_lo), hi = t_hi -
    margin (t_hi - t_lo).  At margin 0, u' = u: the reference's step
    t_lo + u (t_hi - t_lo) (art/sampling.py:877-962)."""
    uf = margin + u * (1.0 - 2.0 * margin)
    return t_lo + (t_hi - t_lo) * uf


def advance(A, points, slack, directions, uniforms, margin=0.0):
    """art/sampling.py:877-962: one move of every chain, in place."""
    Ad = directions @ A.T
    with np.errstate(divide="ignore", invalid="ignore"):
        ratios = np.where(Ad != 0.0, slack / Ad, 0.0)
    t_low = np.max(np.where(Ad < 0.0, ratios, -np.inf), axis=1)
    t_high = np.min(np.where(Ad > 0.0, ratios, np.inf), axis=1)
    step = chord_point(t_low, t_high, uniforms, margin)
    points += directions * step[:, None]
    slack -= Ad * step[:, None]


def shaping_matrix(eigenvectors, eigenvalues, floor_ratio):
    """V diag(sqrt(lambda_floored) / max) (art/sampling.py:496-537, 107-127)."""
    values = np.asarray(eigenvalues, dtype=float)
    floored = np.maximum(values, floor_ratio * float(np.mean(values)))
    s = np.sqrt(floored)
    s /= np.max(s)
    return np.asarray(eigenvectors, dtype=float) * s[None, :]


def run_chains(A, b, starts, n, first, stride, draw, margin=0.0):
    """Generic schedule: chain c makes first + (K_c - 1) * stride moves and
    records after moves first, first + stride, ...; K_c = n // C + (c < n % C);
    record k of chain c is sample k * C + c.  Requires 1 <= C <= n.

    draw(step, live) -> (directions (len(live), d), uniforms (len(live),)) for
    the chains `live` (a prefix 0..L-1) that make move `step` (0-based).
    """
    A, b = np.asarray(A, float), np.asarray(b, float)
    points = np.array(starts, dtype=float)
    C = points.shape[0]
    if not 1 <= C <= n:
        raise ValueError("need 1 <= chains <= n")
    full, rem = divmod(n, C)
    moves = first + (full + (np.arange(C) < rem) - 1) * stride
    slack = b[None, :] - points @ A.T
    out = np.empty((n, A.shape[1]))
    for step in range(int(moves.max())):
        live = np.flatnonzero(moves > step)
        dirs, us = draw(step, live)
        p, s = points[live], slack[live]
        advance(A, p, s, np.array(dirs, dtype=float), np.asarray(us, dtype=float), margin)
        points[live], slack[live] = p, s
        done = step + 1
        if done >= first and (done - first) % stride == 0:
            k = (done - first) // stride
            out[k * C + live] = points[live]
    return out


def batched_sample(A, b, starts, n, burn_in, thinning, directions, uniforms,
                   tail_directions=None, tail_uniforms=None, margin=0.0):
    """BatchedHitAndRunSampler.sample with given arrays (art/sampling.py:618-803):
    directions (burn_in + full*stride, C, d), uniforms (same, C); tail arrays
    (stride, rem, d) / (stride, rem) for the first rem = n % C chains."""
    stride = max(1, thinning)
    n_full = burn_in + (n // np.asarray(starts).shape[0]) * stride

    def draw(step, live):
        if step < n_full:
            return directions[step][live], uniforms[step][live]
        return tail_directions[step - n_full][live], tail_uniforms[step - n_full][live]

    return run_chains(A, b, starts, n, burn_in + stride, stride, draw, margin)


def scalar_sample(A, b, x0, n, burn_in, thinning, directions, uniforms, margin=0.0):
    """HitAndRunSampler.sample with given arrays (art/sampling.py:445-494):
    directions (burn_in + n*thinning, d) normalized before use; uniforms
    (same,) mapped to t_lo + u (t_hi - t_lo)."""
    x = np.array(x0, dtype=float)
    out = []
    for step in range(burn_in + n * thinning):
        d = np.asarray(directions[step], dtype=float)
        nd = np.linalg.norm(d)
        if nd == 0.0:
            continue
        d = d / nd
        lo, hi = step_interval(A, b, x, d)
        x = x + chord_point(lo, hi, uniforms[step], margin) * d
        if step >= burn_in and (step - burn_in) % thinning == 0:
            out.append(x.copy())
    return np.asarray(out)


def box_region(lows, highs):
    """[I; -I] x <= [high; -low] (art/domain.py:94-100)."""
    d = len(lows)
    A = np.vstack([np.eye(d), -np.eye(d)])
    return A, np.concatenate([np.asarray(highs, float), -np.asarray(lows, float)])
