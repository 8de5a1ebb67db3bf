"""Block orthogonalization on the MI355X (BVOrthogonalize with block types gs,
chol, tsqr, tsqrchol and svqb), BVDot and BVMatProject on BVHIPKS bases,
against the oracle and numpy.  The reference's test11 flow (leading columns
first, then the active ones; src/sys/classes/bv/tests/test11.c) on its fill
and at lengths the device handles with many workgroups, ragged row counts
included; test11's TSQR case at n=180, k=7 (output/test11_12.out); a panel
too ill-conditioned for Cholesky QR; constraint columns."""
import numpy as np
import pytest

from slepc_amd import matgen as M

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
BLOCKS = {"gs": 0, "chol": 1, "tsqr": 2, "tsqrchol": 3, "svqb": 4}


@pytest.fixture(scope="module")
def E():
    from slepc_amd import eps as E
    E.initialize(0)
    return E


def fill_test11(n, k):
    X = np.zeros((n, k), order="F")
    for j in range(k):
        for i in range(n // 2 + 1):
            if i + j < n:
                X[i + j, j] = (3.0 * i + j - 2) / (2 * (i + j + 1))
    return X


def _device_flow(E, X, l, block):
    n, k = X.shape
    bv = E.BV(n, k)
    bv.set_columns(X)
    bv.set_block(block)
    R = np.zeros((k, k), order="F")
    if l:
        bv.set_active(0, l)
        R1 = bv.orthogonalize()
        R[:, :l] = R1[:, :l]
    bv.set_active(l, k)
    R2 = bv.orthogonalize()
    R[:, l:] = R2[:, l:]
    bv.set_active(0, k)
    Q = bv.get_columns()
    G = bv.dot(bv)
    bv.destroy()
    return Q, R, G


def _match_signs(Q, R, Qo):
    """SVQB's columns are eigenvectors of the scaled Gram matrix, whose signs dsyev
    fixes from rounding: flip Q's columns and R's rows to the oracle's."""
    s = np.sign(np.einsum("ij,ij->j", Q, Qo))
    s[s == 0] = 1.0
    return Q * s, (R.T * s).T


def _positive_diagonal(Q, R):
    """Householder QR (TSQR, TSQRCHOL) fixes the sign of each row of R by its own
    rule, which a tree of factorizations need not share with one LAPACK call:
    compare with R's diagonal made positive."""
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return Q * s, (R.T * s).T


def _normalize(name, Q, R, Qo, Ro):
    if name == "svqb":
        Q, R = _match_signs(Q, R, Qo)
    elif name in ("tsqr", "tsqrchol"):
        Q, R = _positive_diagonal(Q, R)
        Qo, Ro = _positive_diagonal(Qo, Ro)
    return Q, R, Qo, Ro


def _oracle_flow(oracle, X, l, block):
    k = X.shape[1]
    Y = X.copy(order="F")
    Rf = np.zeros((k, k), order="F")
    if l:
        Y1 = np.asfortranarray(Y[:, :l])
        Rf[:l, :l] = oracle.bv_orthogonalize(Y1, 0, block)
        Y[:, :l] = Y1
    R = oracle.bv_orthogonalize(Y, l, block)
    Rf[:, l:] = R[:, l:]
    return Y, Rf


@pytest.mark.parametrize("name", list(BLOCKS))
def test_test11_on_device(E, oracle, name):
    n, l, k = 20, 2, 8
    X = fill_test11(n, k)
    Q, R, G = _device_flow(E, X, l, BLOCKS[name])
    I = np.eye(k)
    print(f"block={name} orth={np.linalg.norm(G - I):.2e} resid={np.linalg.norm(X - Q @ R):.2e}", flush=True)
    assert np.linalg.norm(G[:l, :l] - I[:l, :l]) < 100 * EPS
    assert np.linalg.norm(G[l:, l:] - I[l:, l:]) < 100 * EPS
    assert np.linalg.norm(G - I) < 100 * EPS
    assert np.linalg.norm(X - Q @ R) < 100 * EPS
    Qo, Ro = _oracle_flow(oracle, X, l, BLOCKS[name])
    Q, R, Qo, Ro = _normalize(name, Q, R, Qo, Ro)
    assert np.abs(Q - Qo).max() < 1e-13
    assert np.abs(R - Ro).max() < 1e-13 * np.abs(Ro).max()


@pytest.mark.parametrize("name", list(BLOCKS))
@pytest.mark.parametrize("n,l,k", [(100003, 3, 16), (1 << 20, 8, 40), (70001, 0, 70)])
def test_large_vs_oracle(E, oracle, name, n, l, k):
    # ragged lengths, widths across one and two 64-column Gram blocks
    rng = np.random.default_rng(n + k)
    X = np.asfortranarray(rng.standard_normal((n, k)))
    print(f"block={name} n={n} l={l} k={k}", flush=True)
    Q, R, G = _device_flow(E, X, l, BLOCKS[name])
    assert np.abs(G - np.eye(k)).max() < 1e-13
    assert np.abs(X - Q @ R).max() < 1e-12 * np.abs(X).max()
    Qo, Ro = _oracle_flow(oracle, X, l, BLOCKS[name])
    Q, R, Qo, Ro = _normalize(name, Q, R, Qo, Ro)
    assert np.abs(Q - Qo).max() < 1e-11
    assert np.abs(R - Ro).max() < 1e-11 * np.abs(Ro).max()


def test_dot_and_mat_project_vs_numpy(E):
    n = 50021
    csr = M.er_laplacian(n, 8.0, 3)
    A = E.Mat(csr)
    rng = np.random.default_rng(1)
    Xa = rng.standard_normal((n, 12))
    Ya = rng.standard_normal((n, 9))
    X, Y = E.BV(n, 12), E.BV(n, 9)
    X.set_columns(Xa)
    Y.set_columns(Ya)
    # BVDot over active columns: entries outside Y(ly:ky)^T X(lx:kx) are left as they were
    X.set_active(2, 12)
    Y.set_active(1, 9)
    M0 = np.full((9, 12), 7.0)
    Md = X.dot(Y, M0)
    ref = np.full((9, 12), 7.0)
    ref[1:, 2:] = Ya[:, 1:].T @ Xa[:, 2:]
    scale = np.abs(Ya).T @ np.abs(Xa)
    assert np.all(np.abs(Md - ref) <= 64 * EPS * np.where(ref == 7.0, 1.0, scale))
    # BVMatProject: M = Y^T A X over all columns
    X.set_active(0, 12)
    Y.set_active(0, 9)
    Mp = X.mat_project(A, Y)
    AX = np.zeros_like(Xa)
    rows = np.repeat(np.arange(n), np.diff(csr[0]))
    for j in range(12):
        np.add.at(AX[:, j], rows, csr[2] * Xa[csr[1], j])
    refp = Ya.T @ AX
    assert np.abs(Mp - refp).max() < 1e-10 * np.abs(refp).max()
    # with A = None the projection is the plain dot product
    Mi = X.mat_project(None, Y)
    assert np.abs(Mi - Ya.T @ Xa).max() < 1e-11 * np.abs(Ya.T @ Xa).max()
    X.destroy()
    Y.destroy()


def test_dot_fp32_basis(E):
    n = 80000
    rng = np.random.default_rng(4)
    Xa = rng.standard_normal((n, 10)).astype(np.float32).astype(np.float64)
    X = E.BV(n, 10, prec=E.F32)
    X.set_columns(Xa)
    G = X.dot(X)
    # FP32 storage, FP64 accumulation: exact products of FP32 values, sums in FP64
    ref = Xa.T @ Xa
    assert np.abs(G - ref).max() < 1e-11 * np.abs(ref).max()
    X.destroy()


def test_test11_12_tsqr_on_device(E):
    # test11 -resid -n 180 -l 0 -k 7 -bv_orthog_block tsqr (output/test11_12.out):
    # "Level of orthogonality of Q < 100*eps", "Residual ||X-Q*R|| < 100*eps"
    X = fill_test11(180, 7)
    Q, R, G = _device_flow(E, X, 0, BLOCKS["tsqr"])
    print(f"orth={np.linalg.norm(G - np.eye(7)):.2e} resid={np.linalg.norm(X - Q @ R):.2e}", flush=True)
    assert np.linalg.norm(G - np.eye(7)) < 100 * EPS
    assert np.linalg.norm(X - Q @ R) < 100 * EPS
    assert np.allclose(np.tril(R, -1), 0.0, atol=0)


@pytest.mark.parametrize("n", [4000, 300007])
def test_tsqr_on_an_ill_conditioned_panel(E, oracle, n):
    # condition number 1e12: TSQR keeps Q orthonormal to rounding and matches the
    # oracle's Householder QR; the Cholesky QR of the same panel does not
    rng = np.random.default_rng(n)
    U, _ = np.linalg.qr(rng.standard_normal((n, 10)))
    W, _ = np.linalg.qr(rng.standard_normal((10, 10)))
    X = np.asfortranarray(U @ np.diag(np.logspace(0, -12, 10)) @ W)
    Q, R, G = _device_flow(E, X, 0, BLOCKS["tsqr"])
    assert np.abs(G - np.eye(10)).max() < 100 * EPS
    assert np.linalg.norm(X - Q @ R) < 100 * EPS * np.linalg.norm(X)
    Qo, Ro = _oracle_flow(oracle, X, 0, BLOCKS["tsqr"])
    Q, R, Qo, Ro = _normalize("tsqr", Q, R, Qo, Ro)
    assert np.abs(R - Ro).max() < 1e-13 * np.abs(Ro).max()
    Qc, Rc, Gc = _device_flow(E, X, 0, BLOCKS["chol"])
    assert np.abs(Gc - np.eye(10)).max() > 1e-6


@pytest.mark.parametrize("name", list(BLOCKS))
@pytest.mark.parametrize("n,l", [(5003, 0), (100003, 2)])
def test_constraint_columns(E, name, n, l):
    # three constraint columns: the active columns come out orthonormal, orthogonal
    # to the constraints, and Q R reproduces their part outside the constraints' span
    rng = np.random.default_rng(n + l)
    k, nc = 9, 3
    Cm = rng.standard_normal((n, nc))
    X = np.asfortranarray(rng.standard_normal((n, k)))
    bv = E.BV(n, k)
    assert bv.insert_constraints(Cm) == nc
    Cq = bv.get_constraints(nc)
    assert np.abs(Cq.T @ Cq - np.eye(nc)).max() < 1e-13
    bv.set_columns(X)
    bv.set_block(BLOCKS[name])
    R = np.zeros((k, k), order="F")
    if l:
        bv.set_active(0, l)
        R[:, :l] = bv.orthogonalize()[:, :l]
    bv.set_active(l, k)
    R[:, l:] = bv.orthogonalize()[:, l:]
    Q = bv.get_columns()
    bv.destroy()
    P = X - Cq @ (Cq.T @ X)
    print(f"block={name} n={n} l={l} |C^T Q|={np.abs(Cq.T @ Q).max():.2e}", flush=True)
    assert np.abs(Cq.T @ Q).max() < 1e-13
    assert np.abs(Q.T @ Q - np.eye(k)).max() < 1e-13
    assert np.abs(P - Q @ R).max() < 1e-12 * np.abs(X).max()


def test_tsqr_needs_an_fp64_basis(E):
    bv = E.BV(1000, 4, prec=E.F32)
    bv.set_columns(np.random.default_rng(0).standard_normal((1000, 4)))
    bv.set_block(E.BV.TSQR)
    with pytest.raises(E.SlepcError):
        bv.orthogonalize()
    bv.destroy()
