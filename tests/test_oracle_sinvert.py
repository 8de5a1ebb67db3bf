"""The oracle's shift-and-invert (STSINVERT with an inner preconditioned CG,
sinvert.c:16-105 and stsolve.c:244-262), pinned to the reference's ex11 golden
(src/eps/tutorials/ex11.c:129-139, output/ex11_1.out: the 10 x 10 grid graph
Laplacian with its constant null vector deflated, four smallest eigenvalues
0.09789, 0.09789, 0.19577, 0.38197, both with plain Krylov-Schur and with
-st_type sinvert -eps_target 0 -st_ksp_type cg -st_pc_type jacobi), and to the
1-D Laplacian's closed-form spectrum near a shift; every pair's residual is
measured with A itself, so a wrong inner solve or back-transform fails it."""
import numpy as np
import pytest

from slepc_amd import matgen as M

EX11 = "0.09789, 0.09789, 0.19577, 0.38197"


def test_ex11_plain_golden(oracle):
    csr = M.grid_graph_laplacian(10)
    r = oracle.eps_solve(csr, nev=4, which=oracle.SMALLEST_REAL, keep=0.2, deflation=np.ones((100, 1)))
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == EX11


@pytest.mark.parametrize("pc_none", [False, True])
def test_ex11_sinvert_golden(oracle, pc_none):
    csr = M.grid_graph_laplacian(10)
    r = oracle.eps_solve(csr, nev=4, ncv=20, which=oracle.TARGET_MAGNITUDE, target=0.0, deflation=np.ones((100, 1)),
                         sinvert=True, pc_none=pc_none)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == EX11
    # the outer test bounds the residual of (A - sigma I)^-1; with the inner rtol of 1e-8 the
    # residual of A itself is larger by about the conditioning of A - sigma I
    assert (r["err"][:4] < 1e-6).all() and r["ksp_its"] > 0


def test_sinvert_closed_form_near_shift(oracle):
    n, sigma = 400, -1e-3
    r = oracle.eps_solve(M.laplacian_1d(n), nev=4, ncv=20, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=sigma,
                         sinvert=True, ksp_rtol=1e-13)
    closed = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    near = closed[np.argsort(np.abs(closed - sigma))][:4]
    assert r["reason"] == 1
    assert np.allclose(r["eigr"][:4], near, rtol=1e-10, atol=0)          # in order of distance to the shift
    # tol bounds the residual of theta = 1/(lambda - sigma); the residual of A is larger by
    # about |lambda - sigma| |A - sigma I| / |lambda| (here up to ~10)
    assert (r["err"][:4] < 1e-8).all()
    # far fewer restarts than without the transformation
    p = oracle.eps_solve(M.laplacian_1d(n), nev=4, ncv=20, tol=1e-10, which=oracle.SMALLEST_REAL)
    assert r["its"] < p["its"]


def test_sinvert_requires_target_which(oracle):
    with pytest.raises(RuntimeError):
        oracle.eps_solve(M.laplacian_1d(50), nev=2, which=oracle.LARGEST_MAGNITUDE, sinvert=True)


# eps test2 "Interior eigenvalues" (src/eps/tests/test2.c, output/test2_1.out): the 1-D
# Laplacian, target 2.1, whose shifted matrix A - 2.1 I is indefinite
TEST2_INTERIOR = {30: "2.10130, 1.89870, 2.30286, 2.50131", 32: "2.09516, 2.28463, 1.90484, 2.47152"}


@pytest.mark.parametrize("n", [30, 32])
def test_interior_indefinite_sinvert_bcgs_golden(oracle, n):
    # shift-and-invert at the target with BiCGStab as the inner solver reaches the
    # reference's interior goldens; CG breaks down on the indefinite shifted matrix
    csr = M.laplacian_1d(n)
    r = oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=2.1, sinvert=True,
                         ksp="bcgs", ksp_rtol=1e-12)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == TEST2_INTERIOR[n]
    closed = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    near = closed[np.argsort(np.abs(closed - 2.1))][:4]
    assert np.allclose(r["eigr"][:4], near, rtol=1e-10, atol=0)
    assert (r["err"][:4] < 1e-9).all() and r["ksp_its"] > 0
    with pytest.raises(RuntimeError):
        oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=2.1, sinvert=True)


def test_bcgs_matches_cg_on_a_definite_shift(oracle):
    # on a positive definite A - sigma I both inner solvers give the same eigenpairs
    csr = M.grid_graph_laplacian(10)
    kw = dict(nev=4, ncv=20, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=-0.05, sinvert=True, ksp_rtol=1e-12)
    a = oracle.eps_solve(csr, ksp="cg", **kw)
    b = oracle.eps_solve(csr, ksp="bcgs", **kw)
    assert np.allclose(a["eigr"][:4], b["eigr"][:4], rtol=1e-10, atol=1e-12)


def _lu_dense(csr, lu):
    n = len(csr[0]) - 1
    L, U = np.eye(n), np.zeros((n, n))
    for i in range(n):
        for q in range(csr[0][i], csr[0][i + 1]):
            j = csr[1][q]
            if j < i:
                L[i, j] = lu[q]
            else:
                U[i, j] = lu[q]
    return L, U


def test_ilu0_factors(oracle):
    # ILU(0)'s defining property: L U equals A - sigma I on A's pattern; on a
    # tridiagonal matrix it has no fill to drop and is the exact LU factorization
    sigma = 2.1
    csr = M.laplacian_2d(9, 7)
    n = len(csr[0]) - 1
    L, U = _lu_dense(csr, oracle.ilu0(csr, sigma))
    A = M.csr_to_dense(*csr) - sigma * np.eye(n)
    P = L @ U
    mask = A != 0
    mask[np.arange(n), np.arange(n)] = True
    assert np.abs((P - A)[mask]).max() < 1e-13
    assert np.abs((P - A)[~mask]).max() > 1e-3            # the fill ILU(0) drops
    csr1 = M.laplacian_1d(40)
    L1, U1 = _lu_dense(csr1, oracle.ilu0(csr1, sigma))
    A1 = M.csr_to_dense(*csr1) - sigma * np.eye(40)
    assert np.abs(L1 @ U1 - A1).max() < 1e-13


@pytest.mark.parametrize("n", [30, 32])
def test_interior_bcgs_ilu_golden(oracle, n):
    # with ILU(0) (the exact LU of the tridiagonal A - 2.1 I) each inner solve takes one iteration
    csr = M.laplacian_1d(n)
    r = oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=2.1, sinvert=True,
                         ksp="bcgs", pc="ilu", ksp_rtol=1e-12)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == TEST2_INTERIOR[n]
    assert 0 < r["ksp_its"] <= r["nmatmult"]           # at most one iteration per application


def test_ilu_fewer_iterations_than_jacobi(oracle):
    # 64 x 63 Laplacian, shift between the 4th and 5th eigenvalues: the same eigenpairs,
    # with ILU(0) in far fewer BiCGStab iterations
    csr = M.laplacian_2d(64, 63)
    n = len(csr[0]) - 1
    closed = np.sort(M.laplacian_2d_eigenvalues(64, 63))
    sigma = 0.5 * (closed[3] + closed[4])
    kw = dict(nev=4, ncv=24, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=sigma, sinvert=True, ksp="bcgs",
              ksp_rtol=1e-12, v0=M.uniform_vector(3, n), want_vectors=False)
    a = oracle.eps_solve(csr, pc="jacobi", **kw)
    b = oracle.eps_solve(csr, pc="ilu", **kw)
    near = closed[np.argsort(np.abs(closed - sigma))][:4]
    assert np.allclose(np.sort(b["eigr"][:4]), np.sort(near), rtol=1e-10, atol=0)
    assert np.allclose(a["eigr"][:4], b["eigr"][:4], rtol=1e-10, atol=0)
    assert b["ksp_its"] < a["ksp_its"] / 4


# The direct inner solve KSPPREONLY + PCLU (the reference's default for STSINVERT,
# stsles.c:47-54): the oracle's dense LU with partial pivoting is pinned to LAPACK's
# dgetrf through scipy (the same pivot rows, the same factors), and its solves to the
# goldens above and to numpy's eigenvalues of a nonsymmetric matrix.

def _dense(csr):
    rp, ci, v = csr
    n = len(rp) - 1
    A = np.zeros((n, n))
    for i in range(n):
        np.add.at(A[i], ci[rp[i]:rp[i + 1]], v[rp[i]:rp[i + 1]])
    return A


def _random_sparse(n, seed, zero_diag_every=3):
    # nonsymmetric, about 8 entries a row, every third diagonal entry zero (pivoting needed)
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        js = np.unique(np.concatenate([rng.integers(0, n, 7), [i]]))
        for j in js:
            rows.append(i); cols.append(j)
            vals.append(0.0 if (j == i and i % zero_diag_every == 0) else rng.standard_normal())
    rowptr = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rowptr, np.array(rows) + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), np.array(cols, dtype=np.int32), np.array(vals)


@pytest.mark.parametrize("n,sigma", [(1, 0.5), (7, 0.0), (130, 0.25), (301, -1.5)])
def test_lu_factors_match_lapack(oracle, n, sigma):
    import scipy.linalg as sl
    csr = _random_sparse(n, seed=n)
    lu, ipiv = oracle.lu(csr, sigma)
    A = _dense(csr) - sigma * np.eye(n)
    lu_ref, piv_ref = sl.lu_factor(A)
    assert (ipiv == piv_ref).all()
    assert np.allclose(lu, lu_ref, rtol=1e-11, atol=1e-12 * np.abs(lu_ref).max())
    # and P A = L U rebuilt from the factors
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    PA = A.copy()
    for k in range(n):
        PA[[k, ipiv[k]]] = PA[[ipiv[k], k]]
    assert np.abs(PA - L @ U).max() <= 1e-13 * max(1.0, np.abs(A).max()) * n


def test_lu_zero_pivot_fails(oracle):
    # diag(1, 2, ..., 10) - 3 I is singular: its third column is zero
    n = 10
    csr = (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.arange(1.0, n + 1))
    with pytest.raises(RuntimeError):
        oracle.lu(csr, 3.0)
    with pytest.raises(RuntimeError):
        oracle.eps_solve(csr, nev=2, which=oracle.TARGET_MAGNITUDE, target=3.0, sinvert=True, ksp="preonly", pc="lu")


def test_ex11_direct_singular_shift_and_near_shift(oracle):
    # ex11's shift 0 is an eigenvalue of the graph Laplacian (the deflated constant vector):
    # the LU of A - 0 I meets a zero pivot, where the reference's ex11 runs CG on the
    # deflated space; a shift just below it factors and finds ex11's eigenvalues
    csr = M.grid_graph_laplacian(10)
    kw = dict(nev=4, ncv=20, which=oracle.TARGET_MAGNITUDE, deflation=np.ones((100, 1)), sinvert=True, ksp="preonly",
              pc="lu")
    with pytest.raises(RuntimeError):
        oracle.eps_solve(csr, target=0.0, **kw)
    r = oracle.eps_solve(csr, target=-0.01, **kw)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == EX11
    assert 0 < r["ksp_its"] <= r["nmatmult"]      # one application of the factors per solve
    # the outer tolerance bounds the residual of theta = 1/(lambda - sigma); for A it grows by
    # about |lambda - sigma| |A - sigma I| / |lambda| (near 10 here)
    assert (r["err"][:4] < 1e-6).all()


@pytest.mark.parametrize("n", [30, 32])
def test_interior_direct_golden(oracle, n):
    csr = M.laplacian_1d(n)
    r = oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=2.1, sinvert=True,
                         ksp="preonly", pc="lu")
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == TEST2_INTERIOR[n]
    closed = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    near = closed[np.argsort(np.abs(closed - 2.1))][:4]
    assert np.allclose(r["eigr"][:4], near, rtol=1e-12, atol=0)


def test_markov_nonsymmetric_direct_vs_numpy(oracle):
    # the Markov model of eps test11 (m = 15, N = 120) near 0.5: A - 0.5 I is nonsymmetric
    # and indefinite, which the LU takes as it is
    csr = M.markov_model(15)
    w = np.linalg.eigvals(_dense(csr))
    near = w[np.argsort(np.abs(w - 0.5))][:4]
    r = oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=0.5, hermitian=False,
                         v0=np.ones(120), sinvert=True, ksp="preonly", pc="lu")
    assert r["reason"] == 1 and r["nconv"] >= 4
    assert np.allclose(np.sort(r["eigr"][:4]), np.sort(near.real), rtol=1e-10, atol=0)
    assert np.allclose(r["eigi"][:4], 0.0, atol=1e-12)
