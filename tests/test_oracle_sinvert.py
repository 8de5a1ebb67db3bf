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
