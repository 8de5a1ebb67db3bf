"""Solver-level parity: EPSSolve (Krylov-Schur) through the C host framework
on the MI355X kernels, against the reference's golden outputs, closed forms,
dense LAPACK and the CPU oracle on the same seeded inputs.  Needs an MI355X.

Tolerances (FP64 path): eigenvalues agree with the oracle / closed form to
1e-10 relative, every returned pair satisfies ||Ax - lambda x|| / |lambda| <= tol
(recomputed by EPSComputeError on the device and by numpy on the host), and
eigenvector subspaces agree by principal angles (cosines >= 1 - 1e-9).
"""
import numpy as np
import pytest

from slepc_amd import matgen as M

pytestmark = pytest.mark.gpu
EPS_MACH = np.finfo(float).eps


@pytest.fixture(scope="module")
def E():
    from slepc_amd import eps as E
    E.initialize(0)
    return E


def _check_pairs(csr, r, tol):
    A = None
    n = len(csr[0]) - 1
    for i in range(r["nconv"]):
        assert r["err"][i] <= tol, f"pair {i}: relative residual {r['err'][i]} > {tol}"
    if n <= 20000 and "X" in r:
        A = M.csr_to_dense(*csr) if n <= 5000 else None
    if A is not None:
        for i in range(r["nconv"]):
            if r["eigi"][i] == 0:
                x = r["X"][:, i]
                lam = r["eigr"][i]
                assert np.linalg.norm(A @ x - lam * x) / abs(lam) <= 10 * tol


def test_ex2_golden(E):
    r = E.solve(M.laplacian_2d(72), nev=4, ncv=20, tol=1e-8)
    assert r["reason"] == E.EPS.CONVERGED_TOL and r["nconv"] >= 4
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == "7.99630, 7.99074, 7.98519, 7.98150"
    _check_pairs(M.laplacian_2d(72), r, 1e-8)


def test_ex4_rdb200_golden(E, oracle):
    # src/eps/tutorials/ex4.c on share/slepc/datafiles/matrices/rdb200.petsc (fixture copy),
    # read with the PETSc-binary MatLoad reader; output/ex4_1.out: the 4 largest-magnitude
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rdb200.petsc")
    A = E.Mat.load(path)
    assert A.n == 200 and A.info()["nnz"] == 1120
    A.destroy()
    csr = E.read_petsc_binary(path)
    v0 = oracle.random_uniform(0, 200)
    r = E.solve(csr, nev=4, ncv=19, tol=1e-8, hermitian=False, v0=v0)
    assert r["reason"] == 1
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == "-35.00752, -34.10419, -33.20131, -32.68111"
    ro = oracle.eps_solve(csr, nev=4, ncv=19, tol=1e-8, hermitian=False, v0=v0)
    assert np.allclose(r["eigr"][:4], ro["eigr"][:4], rtol=1e-10)
    _check_pairs(csr, r, 1e-8)


@pytest.mark.parametrize("lock", [True, False])
def test_eps_test2_golden(E, lock):
    tol = max(1000 * EPS_MACH, 1e-9)
    csr = M.laplacian_1d(30)
    r = E.solve(csr, nev=4, tol=tol, which=E.EPS.LARGEST_REAL, lock=lock)
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == "3.98974, 3.95906, 3.90828, 3.83792"
    r = E.solve(csr, nev=4, tol=tol, which=E.EPS.SMALLEST_REAL, lock=lock)
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == "0.01026, 0.04094, 0.09172, 0.16208"
    _check_pairs(csr, r, 5 * tol)


def test_c1_vs_oracle_and_closed_form(E, oracle):
    # 100x100 5-point Laplacian, nev 10, ncv 32, tol 1e-8, start vector U(-1,1) seed 0
    csr = M.laplacian_2d(100)
    v0 = oracle.random_uniform(0, 10000)
    r = E.solve(csr, nev=10, ncv=32, tol=1e-8, v0=v0)
    ro = oracle.eps_solve(csr, nev=10, ncv=32, tol=1e-8, v0=v0)
    assert r["reason"] == 1 and r["nconv"] >= 10
    closed = M.laplacian_2d_eigenvalues(100)
    for v in r["eigr"]:
        assert np.min(np.abs(closed - v)) / v < 1e-12
    # degenerate pairs may be found in either order of discovery: compare distinct values
    d_gpu = np.unique(np.round(r["eigr"][:10], 10))
    d_ora = np.unique(np.round(ro["eigr"][:10], 10))
    k = min(len(d_gpu), len(d_ora))
    assert np.allclose(np.sort(d_gpu)[::-1][:k], np.sort(d_ora)[::-1][:k], rtol=1e-10)
    _check_pairs(csr, r, 1e-8)


def test_varcoef3d_vs_oracle(E, oracle):
    # the C2 operator family at a reduced grid: 24^3, lognormal faces, nev 16 ncv 48 tol 1e-10
    csr = M.varcoef_laplacian_3d(24, seed=1)
    n = len(csr[0]) - 1
    v0 = oracle.random_uniform(7, n)
    r = E.solve(csr, nev=16, ncv=48, tol=1e-10, v0=v0)
    ro = oracle.eps_solve(csr, nev=16, ncv=48, tol=1e-10, v0=v0)
    assert r["reason"] == 1 and ro["reason"] == 1
    assert np.allclose(r["eigr"][:16], ro["eigr"][:16], rtol=1e-10, atol=0)
    assert abs(r["its"] - ro["its"]) <= 1
    assert (r["err"][:16] <= 1e-10).all()
    s = np.linalg.svd(ro["X"][:, :16].T @ r["X"][:, :16], compute_uv=False)
    assert s.min() > 1 - 1e-9
    # the oracle's residuals agree with the device EPSComputeError
    assert (ro["err"][:16] <= 1e-10).all()


@pytest.mark.parametrize("seed", [1, 2])
def test_symmetric_vs_dense(E, seed):
    csr = M.random_sparse_symmetric(600, 0.02, seed)
    A = M.csr_to_dense(*csr)
    w, U = np.linalg.eigh(A)
    idx = np.argsort(-np.abs(w))[:8]
    r = E.solve(csr, nev=8, ncv=40, tol=1e-12)
    assert r["reason"] == 1
    assert np.allclose(r["eigr"][:8], w[idx], rtol=1e-10, atol=0)
    s = np.linalg.svd(U[:, idx].T @ r["X"][:, :8], compute_uv=False)
    assert s.min() > 1 - 1e-9
    _check_pairs(csr, r, 1e-12)


@pytest.mark.parametrize("which", ["LR", "LM"])
def test_nonsymmetric_vs_dense_and_oracle(E, oracle, which):
    csr = M.random_sparse_nonsymmetric(500, 0.03, 3)
    A = M.csr_to_dense(*csr)
    ev = np.linalg.eigvals(A)
    w = E.EPS.LARGEST_REAL if which == "LR" else E.EPS.LARGEST_MAGNITUDE
    r = E.solve(csr, nev=6, ncv=30, tol=1e-11, which=w, hermitian=False)
    ro = oracle.eps_solve(csr, nev=6, ncv=30, tol=1e-11, which=w, hermitian=False)
    assert r["reason"] == 1
    got = r["eigr"] + 1j * r["eigi"]
    for z in got[:6]:
        assert np.min(np.abs(ev - z)) / abs(z) < 1e-10
    assert np.allclose(r["eigr"][:6], ro["eigr"][:6], rtol=1e-10)
    assert np.allclose(np.abs(r["eigi"][:6]), np.abs(ro["eigi"][:6]), rtol=1e-9, atol=1e-12)
    assert (r["err"] <= 1e-11).all()


def test_complex_pairs_residual(E):
    # rotation-dominated block matrix with well separated complex pairs
    n = 400
    rng = np.random.default_rng(4)
    D = np.zeros((n, n))
    for b in range(n // 2):
        a, c = 1.0 + 3.0 * b / n, 0.5 + b / n
        D[2 * b:2 * b + 2, 2 * b:2 * b + 2] = [[a, c], [-c, a]]
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ D @ Q.T
    csr = M.dense_to_csr(A)
    r = E.solve(csr, nev=6, ncv=40, tol=1e-12, which=E.EPS.LARGEST_MAGNITUDE, hermitian=False)
    assert r["reason"] == 1
    assert (r["eigi"][:6] != 0).all()
    assert (r["err"][:r["nconv"]] <= 1e-12).all()
    ev = np.linalg.eigvals(A)
    for z in (r["eigr"] + 1j * r["eigi"])[:6]:
        assert np.min(np.abs(ev - z)) / abs(z) < 1e-10


def test_deflation_space(E):
    n = 60
    rowptr, col, val = (a.copy() for a in M.laplacian_1d(n))
    val[rowptr[0]] = 1.0
    val[rowptr[n - 1] + 1] = 1.0
    r = E.solve((rowptr, col, val), nev=3, ncv=20, tol=1e-10, which=E.EPS.SMALLEST_REAL,
                deflation=np.ones((n, 1)))
    closed = 2 - 2 * np.cos(np.arange(1, 4) * np.pi / n)
    assert np.allclose(r["eigr"][:3], closed, rtol=1e-9)


def test_breakdown_invariant_subspace(E):
    # a diagonal matrix with 6 distinct values: the Krylov space is exhausted after 6 steps
    n = 3000
    d = np.zeros(n)
    d[:6] = [9.0, 7.0, 5.0, 3.0, 2.0, 1.0]
    d[6:] = 0.5
    csr = M.dense_to_csr(np.diag(d)) if n <= 3000 else None
    v0 = np.zeros(n)
    v0[:6] = 1.0
    r = E.solve(csr, nev=4, ncv=20, tol=1e-10, v0=v0)
    assert r["reason"] == 1 and r["nconv"] >= 4
    assert np.allclose(r["eigr"][:4], [9.0, 7.0, 5.0, 3.0], rtol=1e-12)


def test_full_basis_safe_path(E):
    # ncv == n triggers the safeguard of the last column (BV_OrthogonalizeColumn_Safe)
    n = 12
    csr = M.laplacian_1d(n)
    r = E.solve(csr, nev=4, ncv=12, tol=1e-10, which=E.EPS.LARGEST_REAL)
    closed = np.sort(2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))[::-1]
    assert r["nconv"] >= 4
    assert np.allclose(r["eigr"][:4], closed[:4], rtol=1e-10)


@pytest.mark.parametrize("mixed", [0.0, 1e-5])
def test_step_graphs_replay_the_same_solve(E, mixed):
    # each Arnoldi step captured once per column as a hipGraph and replayed in the
    # later restarts runs the same kernels with the same arguments: the solve is
    # identical to the one with plain launches (restarts, passes, eigenvalues to
    # the last bit), FP64 and mixed precision (FP32 basis phase) alike
    csr = M.laplacian_2d(100)
    v0 = M.uniform_vector(0, 10000)
    a = E.solve(csr, nev=10, ncv=32, tol=1e-10, v0=v0, want_vectors=False, mixed_tol=mixed)
    b = E.solve(csr, nev=10, ncv=32, tol=1e-10, v0=v0, want_vectors=False, mixed_tol=mixed, step_graphs=True)
    assert a["its"] > 2 and a["its"] == b["its"] and a["gs_passes"] == b["gs_passes"]
    assert np.array_equal(a["eigr"], b["eigr"])
    assert (b["err"][:10] <= 1e-10).all()
    # the same solver object given another matrix of the same shape (2A) must not
    # replay the first matrix's graphs: its eigenvalues are exactly twice a's
    A1, A2 = E.Mat(csr), E.Mat((csr[0], csr[1], 2.0 * csr[2]))
    eps = E.EPS()
    eps.setProblemType(E.EPS.HEP)
    eps.setDimensions(10, 32)
    eps.setTolerances(1e-10)
    eps.setWhichEigenpairs(E.EPS.LARGEST_MAGNITUDE)
    if mixed > 0:
        eps.setMixedPrecision(mixed)
    vals = []
    for A in (A1, A2):
        eps.setOperators(A)
        eps.setStepGraphs(True)
        eps.setInitialSpace([v0])
        eps.solve()
        vals.append(np.array([np.real(eps.getEigenvalue(i)) for i in range(10)]))
    eps.destroy()
    assert np.array_equal(vals[0], a["eigr"][:10])
    assert np.allclose(vals[1], 2.0 * a["eigr"][:10], rtol=1e-12, atol=0)
