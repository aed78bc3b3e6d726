"""FP64 oracle for the soft-oblique split optimizer (hot-path pieces a4-a6).

Test infrastructure only (see ``oracle/__init__.py``).

Restates, for one node and one start theta0:
* soft objective, loss and gradient -- art/objectives.py:49-51, 132-177, 314-340;
* tangent projection -- art/splitters.py:71-98;
* Armijo backtracking with frozen child models and AdaptiveAlpha --
  art/optimizers.py:20-121, art/splitters.py:365-601 (refit_during_line_search=False);
* stop reasons, last-feasible min-side fallback and the hard refit --
  art/splitters.py:400-690;
* the annealed fixed-step mode (temperature T_k = T0 (T1/T0)^(k/(K-1)), step
  eta along the negative projected gradient, unit-norm retraction).  The
  reference has no annealed mode; its pieces are the reference's objective,
  gradient and projections, and the mode is pinned by the schedule's known
  answers (T_0 = T0, T_(K-1) = T1, t1 = t0/16 over five iterations giving
  t0 2^-k exactly), by step-by-step composition of the pinned gradient and
  projections on the reference's seed-801 fixture, and by planted mistakes
  (exponent k/K, a step before the tangent projection, a dropped or an extra
  iteration, an ascent step) that these checks must reject;
* the device random start (art/splitters.py:692-700 with the device's own
  stream): FP64 Box-Muller from Philox4x32-10 words, w = normal / |normal|
  (w_0 = 1 when the norm is <= 1e-12, art/_cuda_batched_optimizer.cu:7-23),
  z = numpy.median(X w).  Pinned by the Philox known answers
  (tests/test_oracle_sampling.py), numpy's median and the Gaussian moments.

Pinned by (tests/test_oracle_split.py): central finite differences of the
soft loss (gradient), the tangent-projection identities (w . p_w = 0 on the
sphere), the select/freeze known answers of
tests/test_cuda_optimizer_kernels.py:208-214 and
tests/test_cuda_batched_optimizer_kernels.py:148-170, the AdaptiveAlpha rules
checked case by case against art/optimizers.py:49-66, and brute force (a dense
grid over split angles and offsets for d=2 finds no better hard split than a
converged run by more than the grid spacing allows).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import ridge as R


def sigmoid(t):
    """art/objectives.py:49-51."""
    t = np.clip(t, -50.0, 50.0)
    return 1.0 / (1.0 + np.exp(-t))


def project_unit_w(theta):
    """art/splitters.py:71-83."""
    theta = np.asarray(theta, dtype=float).reshape(-1)
    w = theta[:-1]
    n = np.linalg.norm(w)
    if n <= 1e-12:
        raise ValueError("Cannot project theta with near-zero w.")
    out = theta.copy()
    out[:-1] = w / n
    return out


def project_unit_w_gradient(theta, grad):
    """art/splitters.py:86-98."""
    theta = project_unit_w(theta)
    g = np.asarray(grad, dtype=float).reshape(-1).copy()
    w = theta[:-1]
    g[:-1] = g[:-1] - float(g[:-1] @ w) * w
    return g


@dataclass
class Settings:
    """The SoftObliqueSplitter fields the optimizer reads (art/splitters.py:214-250)."""

    temperature: float = 0.05
    ridge: float = 1e-6
    weight_floor: float = 1e-12
    max_iters: int = 100
    grad_atol: float = 1e-8
    grad_rtol: float = 1e-5
    min_side_points: int = 1
    min_side_fraction: float = 0.05
    min_side_fallback: bool = False
    alpha0: float = 1.0
    rho: float = 0.5
    armijo_c: float = 1e-4
    max_backtracks: int = 25
    adaptive_alpha: bool = True
    alpha_min: float = 1e-12
    alpha_max: float = 1e3
    alpha_grow: float = 10.0
    alpha_recovery: float = 10.0
    heavy_backtrack_threshold: int = 8
    max_line_search_failures: int = 3
    degree: int = 1          # 1: affine [x, 1]; 2: quadratic [1, x, x_i x_j]
    bias_last: bool = True   # the affine model's column order

    def required_side_points(self, n):
        """art/splitters.py:750-752."""
        return max(int(self.min_side_points), int(np.ceil(self.min_side_fraction * n)))


def design(X, s: Settings):
    if s.degree == 1:
        return R.features(X, 1, True, s.bias_last)
    return R.features(X, 2, True, False)


@dataclass
class SoftEval:
    theta: np.ndarray
    pi: np.ndarray
    beta: np.ndarray          # (2, p) left, right
    lam: np.ndarray           # (2,)
    rl: np.ndarray
    rr: np.ndarray
    loss: float = None
    grad: np.ndarray = None


def fit_sides(phi, y, pi, s: Settings, floor=None):
    """Left/right weighted ridge fits (art/objectives.py:232-311)."""
    floor = s.weight_floor if floor is None else floor
    beta = np.empty((2, phi.shape[1]))
    lam = np.empty(2)
    for side, q in enumerate((1.0 - pi, pi)):
        beta[side], lam[side] = R.solve_weighted_ridge(phi, y, q, s.ridge, floor)
    return beta, lam


def evaluate(X, y, phi, theta, T, s: Settings, reference: SoftEval | None = None):
    """art/objectives.py:132-177: gates, then refit unless a reference is given."""
    w, z = theta[:-1], float(theta[-1])
    pi = sigmoid((X @ w - z) / T)
    if reference is None:
        beta, lam = fit_sides(phi, y, pi, s)
        rl, rr = y - phi @ beta[0], y - phi @ beta[1]
    else:
        beta, lam, rl, rr = reference.beta, reference.lam, reference.rl, reference.rr
    return SoftEval(theta.copy(), pi, beta, lam, rl, rr)


def loss_of(e: SoftEval):
    """art/objectives.py:314-323."""
    return float(np.mean((1.0 - e.pi) * e.rl ** 2 + e.pi * e.rr ** 2))


def grad_of(X, e: SoftEval, T):
    """art/objectives.py:325-340."""
    n = X.shape[0]
    coeff = (e.rr ** 2 - e.rl ** 2) * e.pi * (1.0 - e.pi) / n
    return np.concatenate([X.T @ coeff / T, [-float(np.sum(coeff)) / T]])


def value_and_grad(X, y, phi, theta, T, s):
    e = evaluate(X, y, phi, theta, T, s)
    e.loss = loss_of(e)
    e.grad = grad_of(X, e, T)
    return e


@dataclass
class AdaptiveAlpha:
    """art/optimizers.py:20-69."""

    alpha: float = 1.0
    alpha_min: float = 1e-12
    alpha_max: float = 1e3
    grow: float = 10.0
    recovery: float = 10.0
    heavy_backtrack_threshold: int = 8

    def __post_init__(self):
        self.alpha = self._clip(self.alpha)

    def update(self, success, step_size, n_backtracks, rho):
        if not success:
            self.alpha = self._clip(step_size * rho)
        elif n_backtracks == 0:
            self.alpha = self._clip(step_size * self.grow)
        elif n_backtracks >= self.heavy_backtrack_threshold:
            self.alpha = self._clip(step_size * self.recovery)
        else:
            self.alpha = self._clip(step_size / rho)
        return self.alpha

    def _clip(self, a):
        return min(max(float(a), self.alpha_min), self.alpha_max)


def armijo(value_fn, candidate_fn, f0, slope, alpha0, rho, c, max_backtracks):
    """art/optimizers.py:72-121: earliest alpha0 rho^k passing Armijo.

    Returns (theta, value, step, success, n_backtracks, candidates_examined).
    """
    alpha = float(alpha0)
    theta = candidate_fn(alpha)
    value = float(value_fn(theta))
    for k in range(max_backtracks + 1):
        if k > 0:
            alpha *= rho
            theta = candidate_fn(alpha)
            value = float(value_fn(theta))
        if value <= f0 + c * alpha * slope:
            return theta, value, alpha, True, k, k + 1
    return theta, value, alpha, False, max_backtracks, max_backtracks + 1


def optimize(X, y, theta0, s: Settings, T=None):
    """One start: art/splitters.py:365-601 with refit_during_line_search=False.

    Returns a dict with the final theta (before the hard refit), histories,
    stop reason, convergence flag, and the last feasible theta/iteration.
    """
    T = s.temperature if T is None else T
    phi = design(X, s)
    n = X.shape[0]
    required = s.required_side_points(n)
    theta = project_unit_w(theta0)
    cur = value_and_grad(X, y, phi, theta, T, s)
    pg = project_unit_w_gradient(theta, cur.grad)
    g0 = float(np.linalg.norm(pg))
    tol = s.grad_atol + s.grad_rtol * g0
    n_right = int(np.count_nonzero(cur.pi >= 0.5))
    feasible = min(n_right, n - n_right) >= required
    last_feasible = theta.copy() if feasible else None
    last_iter = 0 if feasible else None
    hist = dict(soft_loss_history=[cur.loss], projected_grad_norm_history=[g0],
                step_size_history=[], backtrack_history=[], line_search_loss_history=[],
                logical_candidates=0)
    ctrl = AdaptiveAlpha(s.alpha0, s.alpha_min, s.alpha_max, s.alpha_grow, s.alpha_recovery,
                         s.heavy_backtrack_threshold)
    converged, reason = False, "max_iters"
    for _ in range(s.max_iters):
        direction = -pg
        slope = float(cur.grad @ direction)
        if hist["projected_grad_norm_history"][-1] <= tol:
            converged, reason = True, "gradient_tolerance"
            break
        if slope >= 0.0:
            reason = "non_descent_direction"
            break
        failures = 0
        ls = None
        while True:
            a0 = ctrl.alpha if s.adaptive_alpha else s.alpha0
            ls = armijo(lambda t: loss_of(evaluate(X, y, phi, t, T, s, reference=cur)),
                        lambda a: project_unit_w(theta + a * direction),
                        cur.loss, slope, a0, s.rho, s.armijo_c, s.max_backtracks)
            hist["logical_candidates"] += ls[5]
            if s.adaptive_alpha:
                ctrl.update(ls[3], ls[2], ls[4], s.rho)
            if ls[3]:
                break
            failures += 1
            if not s.adaptive_alpha:
                reason = "line_search_failed"
                break
            if failures >= s.max_line_search_failures:
                reason = "max_line_search_failures"
                break
            if ctrl.alpha <= s.alpha_min:
                reason = "alpha_min_reached"
                break
        if not ls[3]:
            break
        theta = ls[0]
        hist["line_search_loss_history"].append(ls[1])
        cur = value_and_grad(X, y, phi, theta, T, s)
        hist["soft_loss_history"].append(cur.loss)
        pg = project_unit_w_gradient(theta, cur.grad)
        hist["projected_grad_norm_history"].append(float(np.linalg.norm(pg)))
        n_right = int(np.count_nonzero(cur.pi >= 0.5))
        if min(n_right, n - n_right) >= required:
            last_feasible = theta.copy()
            last_iter = len(hist["soft_loss_history"]) - 1
        hist["step_size_history"].append(ls[2])
        hist["backtrack_history"].append(ls[4])
    n_right = int(np.count_nonzero(cur.pi >= 0.5))
    fallback = False
    if min(n_right, n - n_right) < required and s.min_side_fallback and last_feasible is not None:
        theta, fallback = last_feasible, True
    return dict(theta=theta, stop_reason=reason, converged=converged,
                n_iters=len(hist["soft_loss_history"]) - 1, initial_grad_norm=g0,
                grad_tolerance=tol, final_alpha=ctrl.alpha if s.adaptive_alpha else s.alpha0,
                min_side_fallback=fallback, fallback_iteration=last_iter if fallback else None,
                last_beta=cur.beta, last_lambda=cur.lam, **hist)


def anneal_schedule(iters, t0, t1):
    """The annealed mode's temperatures: geometric from T_0 = t0 to
    T_(iters-1) = t1, T_k = t0 (t1/t0)^(k/(iters-1)) (T_0 = t0 for one
    iteration)."""
    k = np.arange(int(iters), dtype=float)
    return t0 * (t1 / t0) ** (k / max(int(iters) - 1, 1))


def annealed(X, y, theta0, s: Settings, iters=200, t0=0.2, t1=0.01, step=0.05):
    """Fixed-step projected gradient with a geometric temperature schedule.

    theta <- retract(theta - step * P(grad_T_k(theta))), T_k from
    anneal_schedule, P the tangent projection of art/splitters.py:86-98,
    retraction w/|w| (art/splitters.py:71-83).  The reference optimizes at
    one fixed temperature; this mode is the split-optimizer configuration's
    annealed schedule built from the reference's objective, gradient and
    projections.  Returns (theta, loss history at the start of each
    iteration).
    """
    phi = design(X, s)
    theta = project_unit_w(theta0)
    losses = []
    for T in anneal_schedule(iters, t0, t1):
        e = value_and_grad(X, y, phi, theta, T, s)
        losses.append(e.loss)
        theta = project_unit_w(theta - step * project_unit_w_gradient(theta, e.grad))
    return theta, np.array(losses)


def hard_result(X, y, theta, s: Settings):
    """art/splitters.py:603-690 (hard_refit=True): route u >= z to the right,
    refit unweighted children, report (w, z, n_left, n_right, loss, left_loss,
    right_loss, beta (2,p), lambda (2,)).  Raises ValueError('min_side_points')."""
    phi = design(X, s)
    theta = project_unit_w(theta)
    w, z = theta[:-1], float(theta[-1])
    right = (X @ w - z) >= 0.0
    n = len(y)
    n_right = int(right.sum())
    n_left = n - n_right
    required = s.required_side_points(n)
    if n_left < required or n_right < required:
        raise ValueError("min_side_points")
    beta = np.empty((2, phi.shape[1]))
    lam = np.empty(2)
    losses = []
    for side, mask in enumerate((~right, right)):
        beta[side], lam[side] = R.solve_weighted_ridge(phi[mask], y[mask], np.ones(mask.sum()),
                                                       s.ridge, 0.0)
        losses.append(float(np.mean((y[mask] - phi[mask] @ beta[side]) ** 2)))
    loss = n_left / n * losses[0] + n_right / n * losses[1]
    return dict(w=w.copy(), z=z, n_left=n_left, n_right=n_right, loss=loss, left_loss=losses[0],
                right_loss=losses[1], beta=beta, lam=lam)


def parent_loss(X, y, s: Settings):
    """art/splitters.py:702-730: unweighted parent fit and its MSE."""
    phi = design(X, s)
    beta, _ = R.solve_weighted_ridge(phi, y, np.ones(len(y)), s.ridge, 0.0)
    return float(np.mean((y - phi @ beta) ** 2))


def initial_theta(X, w):
    """art/splitters.py:692-700 for a given Gaussian draw w: unit w (w[0]=1 on a
    tiny norm), z = median(X w)."""
    w = np.array(w, dtype=float)
    nrm = np.linalg.norm(w)
    if nrm <= 1e-12:
        w[0], nrm = 1.0, 1.0
    w = w / nrm
    return np.concatenate([w, [float(np.median(X @ w))]])


def device_start(X, seed, q, counter, key=None):
    """Random start of problem q of an art_split_init call (see module doc).

    Pair i uses Philox4x32-10 counter (i, q, counter_lo, counter_hi), key
    (seed_lo, seed_hi); a keyed start (key given) uses counter (i, 0, 0, 0)
    and key (key_lo, key_hi) instead.  u1 = ((w0 >> 5) 2^26 + (w1 >> 6) + 1) 2^-53,
    u2 = ((w2 >> 5) 2^26 + (w3 >> 6)) 2^-53, normals r cos(2 pi u2),
    r sin(2 pi u2), r = sqrt(-2 ln u1).
    """
    from .sampling import philox4x32_10

    X = np.asarray(X, dtype=float)
    d = X.shape[1]
    pairs = (d + 1) // 2
    if key is not None:
        seed, q, counter = int(key), 0, 0
    ctr = np.array([[i, q, counter & 0xFFFFFFFF, (counter >> 32) & 0xFFFFFFFF]
                    for i in range(pairs)], dtype=np.uint64).astype(np.uint32)
    key = np.tile(np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF],
                           dtype=np.uint64).astype(np.uint32), (pairs, 1))
    w4 = philox4x32_10(ctr, key).astype(np.uint64)
    u1 = ((w4[:, 0] >> np.uint64(5)).astype(float) * 2.0 ** 26
          + (w4[:, 1] >> np.uint64(6)).astype(float) + 1.0) * 2.0 ** -53
    u2 = ((w4[:, 2] >> np.uint64(5)).astype(float) * 2.0 ** 26
          + (w4[:, 3] >> np.uint64(6)).astype(float)) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    normals = np.empty(2 * pairs)
    normals[0::2] = r * np.cos(2.0 * np.pi * u2)
    normals[1::2] = r * np.sin(2.0 * np.pi * u2)
    w = normals[:d].copy()
    nrm = np.linalg.norm(w)
    if nrm <= 1e-12:
        w[0] = 1.0
    else:
        w /= nrm
    return np.concatenate([w, [float(np.median(X @ w))]])
