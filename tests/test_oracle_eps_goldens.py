"""More of the reference's EPS goldens on the oracle: test10 (grid graph
Laplacian 10 x 11 with its constant null vector deflated, four smallest;
src/eps/tests/test10.c, output/test10_1.out), test20 (1-D Laplacian n=18,
smallest eigenvalue, solved again with ncv + 2; output/test20_1.out) and
test24 (1-D Laplacian n=30 with its two lowest exact eigenvectors as
constraints; output/test24_1.out, where the reference runs LOBPCG: the
eigenvalues are the problem's, so Krylov-Schur must print the same)."""
import numpy as np

from slepc_amd import matgen as M

EPS = np.finfo(float).eps


def test_test10_fiedler(oracle):
    csr = M.grid_graph_laplacian(10, 11)
    r = oracle.eps_solve(csr, nev=4, which=oracle.SMALLEST_REAL, max_it=500, deflation=np.ones((110, 1)))
    assert r["reason"] == 1
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == "0.08101, 0.09789, 0.17890, 0.31749"


def test_test20_two_ncv(oracle):
    tol = max(1000 * EPS, 1e-9)
    csr = M.laplacian_1d(18)
    a = oracle.eps_solve(csr, nev=1, tol=tol, which=oracle.SMALLEST_REAL)   # ncv = min(n, max(2 nev, nev + 15)) = 16
    b = oracle.eps_solve(csr, nev=1, ncv=18, tol=tol, which=oracle.SMALLEST_REAL)
    assert "%.5f" % a["eigr"][0] == "0.02728" and "%.5f" % b["eigr"][0] == "0.02728"


def _test24_constraints(n=30, ncon=2):
    alpha, beta = np.pi / (n + 1), np.sqrt(2.0 / (n + 1))
    j = np.arange(1, n + 1)
    return np.column_stack([np.sin(alpha * j * (i + 1)) * beta for i in range(ncon)])


def test_test24_constraints(oracle):
    r = oracle.eps_solve(M.laplacian_1d(30), nev=4, tol=1e-8, max_it=1200, which=oracle.SMALLEST_REAL,
                         deflation=_test24_constraints())
    assert r["reason"] == 1
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == "0.09172, 0.16208, 0.25131, 0.35847"
