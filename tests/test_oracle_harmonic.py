"""The oracle's harmonic Krylov-Schur (EPS_HARMONIC, krylovschur.c:270-319 with
DSTranslateHarmonic_NHEP, dsnhep.c:466-535), pinned to the reference's eps
test2 "Interior eigenvalues" goldens (src/eps/tests/test2.c:80-95, outputs
test2_1.out for n=30 and test2_3.out for n=32: target 2.1, TARGET_MAGNITUDE,
harmonic extraction) and to what the mathematics fixes: the translated
projected matrix has the harmonic Ritz values of the Krylov space, i.e. the
eigenvalues theta of W^T W y = (theta - tau) W^T V y with W = (A - tau I) V,
computed here from an explicit Arnoldi basis with dense LAPACK; full solves
return the eigenvalues of the dense matrix nearest the target."""
import numpy as np
import pytest
import scipy.linalg as sl

from slepc_amd import matgen as M

EPS = np.finfo(float).eps


@pytest.mark.parametrize("n,golden", [(30, "2.10130, 1.89870, 2.30286, 2.50131"),
                                      (32, "2.09516, 2.28463, 1.90484, 2.47152")])
@pytest.mark.parametrize("lock", [True, False])
def test_eps_test2_interior_golden(oracle, n, golden, lock):
    tol = max(1000 * EPS, 1e-9)
    r = oracle.eps_solve(M.laplacian_1d(n), nev=4, tol=tol, which=oracle.TARGET_MAGNITUDE, target=2.1,
                         extraction=oracle.HARMONIC, lock=lock)
    assert r["reason"] == 1
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == golden
    assert (r["err"][:4] < 5 * tol).all()


def _arnoldi(A, v, m):
    n = A.shape[0]
    V = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    V[:, 0] = v / np.linalg.norm(v)
    for j in range(m):
        w = A @ V[:, j]
        for _ in range(2):
            h = V[:, :j + 1].T @ w
            w -= V[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
    return V, H


@pytest.mark.parametrize("tau", [0.3, -1.7])
def test_translation_gives_harmonic_ritz_values(oracle, tau):
    n, m = 200, 20
    A = M.csr_to_dense(*M.random_sparse_nonsymmetric(n, 0.05, 3))
    V, H = _arnoldi(A, np.random.default_rng(0).standard_normal(n), m)
    Ht, g, gamma = oracle.ds_translate_harmonic(H, tau, H[m, m - 1])
    theta = np.linalg.eigvals(Ht)
    W = (A - tau * np.eye(n)) @ V[:, :m]
    ref = tau + sl.eigvals(W.T @ W, W.T @ V[:, :m])
    scale = np.abs(ref).max()
    for t in theta:                                   # conjugate pairs may come in either order
        assert np.min(np.abs(ref - t)) < 1e-10 * scale
    assert np.isclose(gamma, np.hypot(1.0, np.linalg.norm(g)))
    # the translation changes only the last column
    assert np.array_equal(Ht[:, :-1], H[:m, :m - 1])


def test_nonsymmetric_interior_vs_dense(oracle):
    csr = M.convection_diffusion_2d(30, 10.0, seed=2)
    tau = 150.0                                       # inside the spectrum
    r = oracle.eps_solve(csr, nev=6, ncv=30, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=tau,
                         extraction=oracle.HARMONIC, hermitian=False)
    assert r["reason"] == 1 and (r["err"][:6] < 1e-9).all()
    ev = np.linalg.eigvals(M.csr_to_dense(*csr))
    ev = ev[np.argsort(np.abs(ev - tau))]
    got = r["eigr"][:6] + 1j * r["eigi"][:6]
    for z in got:
        assert np.min(np.abs(ev[:8] - z)) < 1e-8 * np.abs(z)
    # nearest to the target first
    assert np.all(np.diff(np.abs(got - tau)) >= -1e-9)


def test_hermitian_interior_closed_form(oracle):
    # 1-D Laplacian, n = 200: simple eigenvalues 2 - 2 cos(k pi / 201); target 1.0 inside
    n, tau = 200, 1.0
    r = oracle.eps_solve(M.laplacian_1d(n), nev=6, ncv=30, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=tau,
                         extraction=oracle.HARMONIC, hermitian=True)
    assert r["reason"] == 1 and (r["err"][:6] < 1e-9).all()
    closed = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    near = closed[np.argsort(np.abs(closed - tau))][:6]
    assert np.allclose(np.sort(r["eigr"][:6]), np.sort(near), rtol=1e-10, atol=0)
    assert np.abs(r["eigi"][:6]).max() == 0.0
    # eigenvectors are returned with unit norm (BVNormalize) and are the sine modes
    X = r["X"][:, :6]
    assert np.allclose(np.linalg.norm(X, axis=0), 1.0, atol=1e-12)
    i = np.arange(1, n + 1)
    for c in range(6):
        k = int(round(np.arccos(1 - r["eigr"][c] / 2) * (n + 1) / np.pi))
        s = np.sin(i * k * np.pi / (n + 1))
        s /= np.linalg.norm(s)
        assert abs(abs(s @ X[:, c]) - 1.0) < 1e-9


def test_harmonic_differs_from_ritz_on_interior_target(oracle):
    # the same interior target with Ritz extraction takes a different trajectory (the
    # translation is not a no-op) and both end at the same eigenvalues
    csr = M.laplacian_1d(60)
    kw = dict(nev=4, ncv=16, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=2.1)
    h = oracle.eps_solve(csr, extraction=oracle.HARMONIC, **kw)
    r = oracle.eps_solve(csr, extraction=oracle.RITZ, **kw)
    assert h["its"] != r["its"]
    assert np.allclose(np.sort(h["eigr"][:4]), np.sort(r["eigr"][:4]), rtol=1e-9)
