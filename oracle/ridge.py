"""FP64 oracle for leaf features, weighted Gram, trace-scaled ridge and solve.

Test infrastructure only (see ``oracle/__init__.py``).

Restates:
* feature column order  -- art/_cuda_features.py:13-29, art/_cuda_features.cu:5-20
* weighted ridge (normal path) -- art/models.py:209-215, 218-343
* batched left/right fit with status codes -- art/_cuda_batched_cholesky.py:198-235,
  273-332 and art/_cuda_batched_cholesky_kernels.cu:15-121

Pinned by (tests/test_oracle_ridge.py):
* sklearn ``PolynomialFeatures`` order and the literal feature KAT of
  tests/test_cuda_features.py:42-50 (library routine + golden vector);
* SciPy ``cho_factor``/``cho_solve`` on random SPD systems (library routine);
* the status known-answer tests of tests/test_cuda_batched_cholesky.py:81-129,
  211-233 (bad gate -> -1, zero column -> pivot 3, empty side -> pivot 1,
  non-finite assembly -> -2, solve failure -> -3);
* the 2x2 Cholesky KAT [[4,1],[1,3]] x = [6,7] -> [1,2] (art/_cuda_cholesky.py:102-106);
* closed form: ridge trace equals ridge * sum(w * phi**2) / p
  (tests/test_cuda_batched_cholesky.py:71-78).
"""

from __future__ import annotations

import numpy as np

# Status codes of the batched fit (art/_cuda_batched_cholesky.py:52-55).
OK, BAD_GATE, NONFINITE, SOLVE_FAILED = 0, -1, -2, -3


def feature_columns(d: int, degree: int = 2, include_bias: bool = True,
                    bias_last: bool = False) -> np.ndarray:
    """Return a (2, p) int32 table of factor indices; -1 means "factor 1".

    Order follows art/_cuda_features.py:13-29: optional bias, the linear
    terms, then x_i*x_j for i <= j with i outermost (sklearn order).  With
    ``bias_last`` the bias moves after the linear terms (affine models).
    """
    if degree not in (1, 2):
        raise ValueError("GPU features support only degree 1 or 2.")
    if bias_last and degree != 1:
        raise ValueError("bias_last is reserved for affine features (degree 1).")
    pairs = [(i, -1) for i in range(d)]
    if degree == 2:
        pairs.extend((i, j) for i in range(d) for j in range(i, d))
    if include_bias:
        pairs.insert(len(pairs) if bias_last else 0, (-1, -1))
    return np.asarray(pairs, dtype=np.int32).T.copy()


def features(X: np.ndarray, degree: int = 2, include_bias: bool = True,
             bias_last: bool = False) -> np.ndarray:
    """Basis expansion of X (..., d) -> (..., p), art/_cuda_features.cu:5-20."""
    X = np.asarray(X, dtype=np.float64)
    a, b = feature_columns(X.shape[-1], degree, include_bias, bias_last)
    ones = np.ones(X.shape[:-1] + (1,))
    ext = np.concatenate([X, ones], axis=-1)  # index -1 selects the trailing 1
    return ext[..., a] * ext[..., b]


def cholesky_lower(a: np.ndarray) -> tuple[np.ndarray, int]:
    """Unblocked lower Cholesky returning (L, info) with LAPACK potrf semantics.

    info = 0 on success, else the one-based order of the first leading minor
    that is not positive definite (pivot <= 0 or NaN).  The reference obtains
    the same code from cusolverDnDpotrfBatched
    (art/_cuda_batched_cholesky.py:315-316).
    """
    a = np.array(a, dtype=np.float64, copy=True)
    p = a.shape[0]
    L = np.zeros_like(a)
    for j in range(p):
        s = a[j, j] - np.dot(L[j, :j], L[j, :j])
        if not s > 0.0:  # catches NaN as well as non-positive pivots
            return L, j + 1
        L[j, j] = np.sqrt(s)
        if j + 1 < p:
            L[j + 1:, j] = (a[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / L[j, j]
    return L, 0


def cholesky_solve(L: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Solve L L^T x = b by forward and back substitution."""
    p = L.shape[0]
    z = np.zeros(p)
    for i in range(p):
        z[i] = (b[i] - L[i, :i] @ z[:i]) / L[i, i]
    x = np.zeros(p)
    for i in range(p - 1, -1, -1):
        x[i] = (z[i] - L[i + 1:, i] @ x[i + 1:]) / L[i, i]
    return x


def weighted_normal_equations(phi, y, w):
    """G = Phi^T diag(w) Phi and h = Phi^T (w y), art/models.py:256-258."""
    phi = np.asarray(phi, dtype=np.float64)
    return phi.T @ (phi * w[:, None]), phi.T @ (y * w)


def scaled_ridge(gram: np.ndarray, ridge: float) -> float:
    """lambda_eff = ridge * trace(G) / p, art/models.py:209-215."""
    return float(ridge) * float(np.trace(gram)) / gram.shape[0]


def solve_weighted_ridge(phi, y, weights, ridge, weight_floor=1e-12):
    """Normal-equation weighted ridge, art/models.py:218-343 (solver="normal").

    Returns (coefficients, ridge_effective).  Raises ``np.linalg.LinAlgError``
    when the Cholesky factorization fails, as the reference does for
    solver="normal" (art/models.py:280-286).
    """
    w = np.maximum(np.asarray(weights, dtype=np.float64), float(weight_floor))
    gram, rhs = weighted_normal_equations(phi, np.asarray(y, np.float64), w)
    lam = scaled_ridge(gram, ridge)
    reg = gram.copy()
    reg.flat[:: gram.shape[0] + 1] += lam
    L, info = cholesky_lower(reg)
    if info:
        raise np.linalg.LinAlgError("Normal-equation Cholesky factorization failed.")
    return cholesky_solve(L, rhs), lam


def solve_system(gram, rhs, ridge):
    """One system through the batched status pipeline.

    Follows regularize_system / protect_system / finish_system in
    art/_cuda_batched_cholesky_kernels.cu:29-105: returns (beta, lambda, status).
    ``lambda`` is 0 when ridge is 0, else ridge * trace / p; status -2 for a
    non-finite trace or lambda, the potrf pivot index on factor failure, -2 for
    a non-finite factor or right-hand side, -3 for a non-finite solve.  Failed
    systems return NaN coefficients.
    """
    p = gram.shape[0]
    trace = float(np.trace(gram))
    lam = 0.0 if ridge == 0.0 else ridge * (trace / p)
    nan = np.full(p, np.nan)
    if not np.isfinite(trace) or not np.isfinite(lam):
        return nan, lam, NONFINITE
    reg = gram.copy()
    reg.flat[:: p + 1] += lam
    L, info = cholesky_lower(reg)
    if info:
        return nan, lam, info
    if not np.all(np.isfinite(L[np.tril_indices(p)])) or not np.all(np.isfinite(rhs)):
        return nan, lam, NONFINITE
    beta = cholesky_solve(L, rhs)
    if not np.all(np.isfinite(beta)):
        return nan, lam, SOLVE_FAILED
    return beta, lam, OK


def batched_gate_fit(phi, y, pi, ridge=0.0, floor=1e-12):
    """Left/right fits for every node: art/_cuda_batched_cholesky.py:273-332.

    phi (K,n,p), y (K,n), pi (K,n) gates.  Side 0 (left) uses
    q = max(1 - pi, floor), side 1 (right) q = max(pi, floor); a gate that is
    non-finite or outside [0, 1] marks both sides of its node -1
    (art/_cuda_batched_cholesky_kernels.cu:15-25).  Returns
    coefficients (K,2,p), status (K,2) int32, ridge_effective (K,2).
    """
    phi, y, pi = (np.asarray(a, dtype=np.float64) for a in (phi, y, pi))
    K, n, p = phi.shape
    coef = np.full((K, 2, p), np.nan)
    status = np.zeros((K, 2), dtype=np.int32)
    lam = np.zeros((K, 2))
    for j in range(K):
        badrow = ~np.isfinite(pi[j]) | (pi[j] < 0.0) | (pi[j] > 1.0)
        g_pi = np.where(badrow, 0.5, pi[j])
        for side in range(2):
            q = np.maximum(1.0 - g_pi if side == 0 else g_pi, floor)
            q[badrow] = 0.0  # a bad gate contributes weight 0 (kernels.cu:21-24)
            g, h = weighted_normal_equations(phi[j], y[j], q)
            coef[j, side], lam[j, side], status[j, side] = solve_system(g, h, ridge)
            if badrow.any():
                # The gate flag overrides every later status (kernels.cu:46-47).
                coef[j, side] = np.nan
                status[j, side] = BAD_GATE
    return coef, status, lam


def leaf_fit(X, y, w, offsets, ridge, floor=1e-12, degree=2, include_bias=True,
             bias_last=False):
    """Variable-length leaves: one weighted ridge fit per CSR row range.

    X (N,d), y (N,), w (N,), offsets (K+1,).  Each leaf j solves the
    normal-path weighted ridge of art/models.py:218-343 on rows
    offsets[j]:offsets[j+1] with features built as in
    art/_cuda_features.py:13-29.  Returns coefficients (K,p), status (K,),
    ridge_effective (K,).
    """
    K = len(offsets) - 1
    phi_all = features(X, degree, include_bias, bias_last)
    p = phi_all.shape[1]
    coef = np.full((K, p), np.nan)
    status = np.zeros(K, dtype=np.int32)
    lam = np.zeros(K)
    for j in range(K):
        s, e = int(offsets[j]), int(offsets[j + 1])
        q = np.maximum(w[s:e], floor)
        g, h = weighted_normal_equations(phi_all[s:e], y[s:e], q)
        coef[j], lam[j], status[j] = solve_system(g, h, ridge)
    return coef, status, lam


def leaf_fit_fast(X, y, w, offsets, ridge, floor=1e-12, degree=2,
                  include_bias=True, bias_last=False):
    """Same arithmetic as :func:`leaf_fit` with LAPACK (SciPy) Cholesky.

    Used as the CPU throughput baseline: per leaf Phi^T W Phi by BLAS,
    trace-scaled ridge, ``cho_factor``/``cho_solve`` -- the calls the
    reference makes in art/models.py:256-343.  Failed leaves get NaN and
    status = 1 (the pivot index is not reported by SciPy).
    """
    from scipy import linalg

    K = len(offsets) - 1
    p = feature_columns(X.shape[1], degree, include_bias, bias_last).shape[1]
    coef = np.full((K, p), np.nan)
    status = np.zeros(K, dtype=np.int32)
    lam = np.zeros(K)
    for j in range(K):
        s, e = int(offsets[j]), int(offsets[j + 1])
        phi = features(X[s:e], degree, include_bias, bias_last)
        q = np.maximum(w[s:e], floor)
        g = phi.T @ (phi * q[:, None])
        h = phi.T @ (y[s:e] * q)
        lam[j] = scaled_ridge(g, ridge)
        g.flat[:: p + 1] += lam[j]
        try:
            coef[j] = linalg.cho_solve(linalg.cho_factor(g, lower=True, check_finite=False),
                                       h, check_finite=False)
        except linalg.LinAlgError:
            status[j] = 1
    return coef, status, lam
