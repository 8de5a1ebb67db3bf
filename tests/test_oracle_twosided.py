"""The oracle's two-sided Krylov-Schur (EPSSetTwoSided, ks-twosided.c:138-241 with
DSNHEPTS, dsnhepts.c), pinned to the reference's goldens that run with
-eps_two_sided {0,1}: ex5 (Markov model, m=15, largest real, locking on and
off; src/eps/tutorials/ex5.c:172-178, output/ex5_1.out) and ex9 (Brusselator,
n=50, output/ex9_1.out); and to what the mathematics fixes: every returned
left vector y satisfies A^T y = lambda y against dense LAPACK's left
eigenvectors, and left and right eigenvectors of distinct eigenvalues are
biorthogonal (y_i^T x_j = 0)."""
import numpy as np
import pytest
import scipy.linalg as sl

from slepc_amd import matgen as M

EX5 = "1.00000, 0.97137, 0.90423, 0.85714"
EX9 = "0.00007+2.13946i, 0.00007-2.13946i, -0.67386+2.52812i, -0.67386-2.52812i"


def _ex5_start(n):
    v0 = np.zeros(n)
    v0[:3] = 1.0                        # ex5.c:62-69
    return v0


@pytest.mark.parametrize("twosided", [False, True])
@pytest.mark.parametrize("lock", [True, False])
def test_ex5_markov_golden(oracle, twosided, lock):
    csr = M.markov_model(15)
    n = len(csr[0]) - 1
    assert n == 120
    r = oracle.eps_solve(csr, nev=4, which=oracle.LARGEST_REAL, hermitian=False, v0=_ex5_start(n),
                         twosided=twosided, lock=lock)
    assert r["reason"] == 1
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == EX5


@pytest.mark.parametrize("twosided", [False, True])
def test_ex9_brusselator_golden(oracle, twosided):
    r = oracle.eps_solve(M.brusselator(50), nev=4, which=oracle.LARGEST_REAL, hermitian=False, twosided=twosided)
    assert r["reason"] == 1
    assert ", ".join("%.5f%+.5fi" % (a, b) for a, b in zip(r["eigr"][:4], r["eigi"][:4])) == EX9


def test_markov_left_vectors_vs_dense(oracle):
    csr = M.markov_model(15)
    n = len(csr[0]) - 1
    A = M.csr_to_dense(*csr)
    r = oracle.eps_solve(csr, nev=4, tol=1e-12, which=oracle.LARGEST_REAL, hermitian=False, v0=_ex5_start(n),
                         twosided=True)
    w, vl = sl.eig(A, left=True, right=False)
    X, Y = r["X"], r["Y"]
    for i in range(4):
        lam = r["eigr"][i]
        y = Y[:, i] / np.linalg.norm(Y[:, i])
        assert np.linalg.norm(A.T @ y - lam * y) < 1e-9
        j = np.argmin(np.abs(w - lam))
        ref = np.real(vl[:, j]) / np.linalg.norm(np.real(vl[:, j]))
        assert abs(abs(ref @ y) - 1.0) < 1e-9
    # A is the transpose of a stochastic matrix: its left eigenvector for 1 is constant
    y0 = Y[:, 0] / Y[0, 0]
    assert np.allclose(y0, 1.0, atol=1e-9)
    # biorthogonality of left and right eigenvectors of distinct eigenvalues
    G = Y.T @ X
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() < 1e-9 * np.abs(np.diag(G)).min()


def test_nonsymmetric_twosided_matches_one_sided(oracle):
    csr = M.convection_diffusion_2d(24, 10.0, seed=2)
    kw = dict(nev=6, ncv=24, tol=1e-10, which=oracle.LARGEST_REAL, hermitian=False)
    a = oracle.eps_solve(csr, **kw)
    b = oracle.eps_solve(csr, twosided=True, **kw)
    za = np.sort_complex(a["eigr"][:6] + 1j * a["eigi"][:6])
    zb = np.sort_complex(b["eigr"][:6] + 1j * b["eigi"][:6])
    assert np.allclose(za, zb, rtol=1e-9, atol=0)
    A = M.csr_to_dense(*csr)
    for i in range(6):
        y = b["Y"][:, i]
        if b["eigi"][i] == 0.0:
            assert np.linalg.norm(A.T @ y - b["eigr"][i] * y) / np.linalg.norm(y) < 1e-8 * abs(b["eigr"][i])
