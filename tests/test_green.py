"""The closed form behind the GPU's truncated Green's-function line solves:
for the normalised periodic system x_i + a (x_{i-1} + x_{i+1}) = d_i
(a = -G/(1+2G), solxi.f90:33-37), the filter with
g_m = C (rho^m + rho^(N-m)) / (1 - rho^N), rho = -2a / (1 + sqrt(1-4a^2)),
C = 1 / (1 + 2 a rho), truncated where rho^(K+1) < tol, reproduces the
dense solve (numpy) to round-off, and the truncation length K is what the
library uses (K <= 12, else cyclic reduction)."""
import numpy as np
import pytest


def green(G, N, tol):
    a = -G / (1 + 2 * G)
    rho = -2 * a / (1 + np.sqrt(1 - 4 * a * a))
    C = 1 / (1 + 2 * a * rho)
    K = 0
    while K < N // 2 - 1 and rho ** (K + 1) > tol:
        K += 1
    g = np.array([C * (rho ** m + rho ** (N - m)) / (1 - rho ** N) for m in range(K + 1)])
    return K, g


@pytest.mark.parametrize("G,N", [(1e-3, 32), (9.4e-3, 256), (0.05, 48), (0.2, 96), (4e-4, 24)])
def test_truncated_green_matches_dense_solve(G, N):
    K, g = green(G, N, 2e-17)
    assert K <= 12 or G > 0.1   # G = 0.2 needs 19 terms: the library solves it by cyclic reduction
    a = -G / (1 + 2 * G)
    A = np.eye(N) + a * (np.roll(np.eye(N), 1, axis=1) + np.roll(np.eye(N), -1, axis=1))
    d = np.random.default_rng(1).normal(size=N)
    x = np.linalg.solve(A, d)
    y = g[0] * d + sum(g[m] * (np.roll(d, m) + np.roll(d, -m)) for m in range(1, K + 1))
    assert np.max(np.abs(x - y)) <= 1e-14 * np.max(np.abs(x))


def test_slow_decay_needs_cyclic_reduction():
    K, _ = green(25.0, 256, 2e-17)  # a -> -1/2: the filter would be long
    assert K > 12
