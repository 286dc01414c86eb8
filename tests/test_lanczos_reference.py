"""The numpy Lanczos checker against what the mathematics fixes: a full
Krylov run reproduces the spectrum of a symmetric matrix, partial runs
interlace it (Cauchy), the basis is orthonormal and V^T A V is the
tridiagonal matrix."""
import numpy as np

from tests.cases import lanczos_reference, tridiagonal_eigen


def sym(n, seed):
    a = np.random.default_rng(seed).standard_normal((n, n))
    return a + a.T


def test_full_run_is_the_spectrum():
    A = sym(60, 1)
    v0 = np.random.default_rng(2).standard_normal(60)
    al, be, V = lanczos_reference(lambda x: A @ x, v0, 60)
    th, _ = tridiagonal_eigen(al, be)
    assert np.allclose(th, np.linalg.eigvalsh(A), atol=1e-9 * np.abs(th).max())


def test_partial_run_interlaces_and_is_orthonormal():
    A = sym(80, 3)
    lam = np.linalg.eigvalsh(A)
    v0 = np.random.default_rng(4).standard_normal(80)
    m = 12
    al, be, V = lanczos_reference(lambda x: A @ x, v0, m)
    th, S = tridiagonal_eigen(al, be)
    eps = 1e-9 * np.abs(lam).max()
    for k in range(m):
        assert th[k] >= lam[k] - eps
        assert th[m - 1 - k] <= lam[80 - 1 - k] + eps
    Q = V[:m]
    assert np.allclose(Q @ Q.T, np.eye(m), atol=1e-12)
    T = np.diag(al) + np.diag(be[:m - 1], 1) + np.diag(be[:m - 1], -1)
    assert np.allclose(Q @ A @ Q.T, T, atol=1e-10 * np.abs(lam).max())
    # residual bound |A y - theta y| = |beta_m s_m|
    Y = S.T @ Q
    for i in range(m):
        r = np.linalg.norm(A @ Y[i] - th[i] * Y[i])
        assert abs(r - abs(be[m - 1] * S[m - 1, i])) <= 1e-8 * np.abs(lam).max()
