"""Block orthogonalization on the MI355X (BVOrthogonalize with block types gs,
chol and svqb), BVDot and BVMatProject on BVHIPKS bases, against the oracle
and numpy.  The reference's test11 flow (leading columns first, then the
active ones; src/sys/classes/bv/tests/test11.c) on its fill and at lengths the
device handles with many workgroups, ragged row counts included."""
import numpy as np
import pytest

from slepc_amd import matgen as M

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
BLOCKS = {"gs": 0, "chol": 1, "svqb": 4}


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
    if name == "svqb":
        Q, R = _match_signs(Q, R, Qo)
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
    if name == "svqb":
        Q, R = _match_signs(Q, R, Qo)
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


def test_tsqr_is_unsupported(E):
    bv = E.BV(1000, 4)
    bv.set_columns(np.random.default_rng(0).standard_normal((1000, 4)))
    bv.set_block(E.BV.TSQR)
    with pytest.raises(E.SlepcError):
        bv.orthogonalize()
    bv.destroy()
