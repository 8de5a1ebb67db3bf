"""The oracle's block orthogonalization (BVOrthogonalize with -bv_orthog_block
gs, chol, tsqr, tsqrchol, svqb) and BVDot, pinned to the reference's test11
(src/sys/classes/bv/tests/test11.c, outputs test11_1.out and test11_6.out:
orthogonality of Q1, Q2, Q and the residual ||X - Q R|| all below 100 eps for
n=20, l=2, k=8; test11_12.out: the residual for n=180, l=0, k=7 with tsqr)
and to what the mathematics fixes: GS and CHOL give the unique QR
factorization with a positive diagonal (numpy's QR after sign normalization),
TSQR and TSQRCHOL the same factorization up to the signs of R's rows (a
Householder QR), also for a panel too ill-conditioned for CHOL, CHOL's R is the Cholesky factor of X^T X, and SVQB's Q^T X D
has orthogonal rows of ascending norm (D = diag(1/||x_i||)), which fails if
the diagonal scaling or the order of the eigenpairs is lost."""
import numpy as np
import pytest

EPS = np.finfo(float).eps
BLOCKS = {"gs": 0, "chol": 1, "tsqr": 2, "tsqrchol": 3, "svqb": 4}


def fill_test11(n, k):
    """Column j holds (3i+j-2)/(2(i+j+1)) at row i+j, i = 0..n/2 (test11.c:75-88)."""
    X = np.zeros((n, k), order="F")
    for j in range(k):
        for i in range(n // 2 + 1):
            if i + j < n:
                X[i + j, j] = (3.0 * i + j - 2) / (2 * (i + j + 1))
    return X


def _test11_flow(oracle, X, l, block):
    """Leading columns first, then the active ones, R assembled as the test does."""
    n, k = X.shape
    Y = X.copy(order="F")
    Rf = np.zeros((k, k), order="F")
    if l:
        Y1 = np.asfortranarray(Y[:, :l])
        R1 = oracle.bv_orthogonalize(Y1, 0, block)
        Y[:, :l] = Y1
        Rf[:l, :l] = R1
    R = oracle.bv_orthogonalize(Y, l, block)
    Rf[:, l:] = R[:, l:]
    return Y, Rf


@pytest.mark.parametrize("name", list(BLOCKS))
def test_test11_golden_properties(oracle, name):
    n, l, k = 20, 2, 8
    X = fill_test11(n, k)
    Y, R = _test11_flow(oracle, X, l, BLOCKS[name])
    M = np.zeros((k, k), order="F")
    oracle.bv_dot(Y, Y, M)
    I = np.eye(k)
    assert np.linalg.norm(M[:l, :l] - I[:l, :l]) < 100 * EPS        # "Level of orthogonality of Q1"
    assert np.linalg.norm(M[l:, l:] - I[l:, l:]) < 100 * EPS        # "... of Q2"
    assert np.linalg.norm(M - I) < 100 * EPS                         # "... of Q"
    assert np.linalg.norm(X[:, :l] - Y[:, :l] @ R[:l, :l]) < 100 * EPS   # "Residual ||X1-Q1*R11||"
    assert np.linalg.norm(X - Y @ R) < 100 * EPS                     # "Residual ||X-Q*R||"


def _positive_diagonal(Q, R):
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return Q * s, (R.T * s).T


@pytest.mark.parametrize("name", ["gs", "chol", "tsqr", "tsqrchol"])
def test_triangular_factor_is_numpy_qr(oracle, name):
    rng = np.random.default_rng(3)
    X = np.asfortranarray(rng.standard_normal((300, 12)))
    Y, R = _test11_flow(oracle, X, 4, BLOCKS[name])
    Y, R = _positive_diagonal(Y, R)     # Householder QR fixes the signs of R's rows its own way
    Qn, Rn = _positive_diagonal(*np.linalg.qr(X))
    assert np.allclose(np.tril(R, -1), 0.0, atol=0)
    assert np.abs(Y - Qn).max() < 1e-12
    assert np.abs(R - Rn).max() < 1e-12 * np.abs(Rn).max()


def test_chol_factor_is_cholesky_of_gram(oracle):
    rng = np.random.default_rng(5)
    X = np.asfortranarray(rng.standard_normal((200, 9)))
    Y = X.copy(order="F")
    R = oracle.bv_orthogonalize(Y, 0, BLOCKS["chol"])
    L = np.linalg.cholesky(X.T @ X)
    assert np.abs(R - L.T).max() < 1e-11 * np.abs(L).max()


def test_svqb_scaled_eigenbasis(oracle):
    rng = np.random.default_rng(7)
    X = np.asfortranarray(rng.standard_normal((400, 10)) * np.logspace(0, 3, 10))   # columns of very different norm
    Y = X.copy(order="F")
    R = oracle.bv_orthogonalize(Y, 0, BLOCKS["svqb"])
    assert np.abs(Y.T @ Y - np.eye(10)).max() < 1e-13
    assert np.abs(X - Y @ R).max() < 1e-12 * np.abs(X).max()
    D = np.diag(1.0 / np.linalg.norm(X, axis=0))
    W = Y.T @ X @ D                               # = Lambda^1/2 U^T: rows orthogonal
    G = W @ W.T
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() < 1e-12 * np.abs(G).max()
    assert np.all(np.diff(np.diag(G)) > 0)        # dsyev order: ascending eigenvalues
    # without the scaling, Y^T X does not have orthogonal rows
    W0 = Y.T @ X
    G0 = W0 @ W0.T
    assert np.abs(G0 - np.diag(np.diag(G0))).max() > 1e-3 * np.abs(G0).max()


def test_bv_dot_leading_columns(oracle):
    rng = np.random.default_rng(9)
    X = np.asfortranarray(rng.standard_normal((50, 7)))
    Y = np.asfortranarray(rng.standard_normal((50, 5)))
    M = np.full((5, 7), 42.0, order="F")
    oracle.bv_dot(X, Y, M, lx=2, ly=1)
    ref = np.full((5, 7), 42.0)
    ref[1:, 2:] = Y[:, 1:].T @ X[:, 2:]
    assert np.abs(M - ref).max() < 1e-13


def test_chol_retry_on_semidefinite_gram(oracle):
    # a numerically dependent column: the first potrf fails, the retry with 50 eps on
    # the diagonal succeeds (bvlapack.c:171-179)
    rng = np.random.default_rng(11)
    X = rng.standard_normal((60, 4))
    X = np.asfortranarray(np.column_stack([X, X[:, 0]]))
    Y = X.copy(order="F")
    R = oracle.bv_orthogonalize(Y, 0, BLOCKS["chol"])
    assert np.isfinite(R).all() and np.isfinite(Y).all()


def test_test11_12_tsqr_residual(oracle):
    # test11 -resid -n 180 -l 0 -k 7 -bv_orthog_block tsqr on 7 processes
    # (output/test11_12.out): the TSQR over 7 row blocks meets both figures; one
    # Householder QR of the whole panel leaves an absolute residual of about
    # 2.5e-14 (||X|| = 33), 8e-16 relative, just above the absolute 100 eps
    X = fill_test11(180, 7)
    Y = X.copy(order="F")
    R = oracle.bv_tsqr_split(Y, 7)
    assert np.linalg.norm(Y.T @ Y - np.eye(7)) < 100 * EPS
    assert np.linalg.norm(X - Y @ R) < 100 * EPS
    Y1, R1 = _test11_flow(oracle, X, 0, BLOCKS["tsqr"])
    assert np.linalg.norm(Y1.T @ Y1 - np.eye(7)) < 100 * EPS
    assert np.linalg.norm(X - Y1 @ R1) < 100 * EPS * np.linalg.norm(X)
    Y, R = _positive_diagonal(Y, R)
    Y1, R1 = _positive_diagonal(Y1, R1)
    assert np.abs(R - R1).max() < 1e-13 * np.abs(R1).max() and np.abs(Y - Y1).max() < 1e-13


def test_tsqr_on_a_panel_too_ill_conditioned_for_chol(oracle):
    # condition number 1e12: Householder QR keeps Q orthonormal to rounding, the
    # Cholesky QR loses about kappa^2 u of it
    rng = np.random.default_rng(5)
    U, _ = np.linalg.qr(rng.standard_normal((400, 10)))
    W, _ = np.linalg.qr(rng.standard_normal((10, 10)))
    X = np.asfortranarray(U @ np.diag(np.logspace(0, -12, 10)) @ W)
    Y = X.copy(order="F")
    R = oracle.bv_orthogonalize(Y, 0, BLOCKS["tsqr"])
    assert np.linalg.norm(Y.T @ Y - np.eye(10)) < 100 * EPS
    assert np.linalg.norm(X - Y @ R) < 100 * EPS * np.linalg.norm(X)
    Yc = X.copy(order="F")
    oracle.bv_orthogonalize(Yc, 0, BLOCKS["chol"])
    assert np.linalg.norm(Yc.T @ Yc - np.eye(10)) > 1e-6


def test_tsqrchol_factor_equals_tsqr_factor(oracle):
    rng = np.random.default_rng(7)
    X = np.asfortranarray(rng.standard_normal((250, 9)))
    Y1, Y2 = X.copy(order="F"), X.copy(order="F")
    R1 = oracle.bv_orthogonalize(Y1, 3, BLOCKS["tsqr"])
    R2 = oracle.bv_orthogonalize(Y2, 3, BLOCKS["tsqrchol"])
    assert np.array_equal(R1, R2)     # the same dgeqrf on the same panel
    assert np.abs(Y1 - Y2).max() < 1e-12
