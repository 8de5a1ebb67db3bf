"""Shift-and-invert on the MI355X (STSINVERT; the inner solves by the device
CG of ks_pcg_solve or BiCGStab of ks_bcgs_solve with Jacobi, none or the
ILU(0) of ks_ilu0_create, through STGetKSP/KSPSetType/KSPSetTolerances/
PCSetType): the reference's ex11 golden
with -st_type sinvert -st_ksp_type cg -st_pc_type jacobi, the device against
the oracle (restarts, inner iterations, eigenvalues), the interior goldens of
eps test2 (target 2.1 inside the spectrum, A - sigma I indefinite) with
BiCGStab, closed forms near a shift at device size, and the error paths (a
non-target 'which', an indefinite shifted matrix for CG -- every kernel exits
normally and the host reports the failed solve)."""
import numpy as np
import pytest

from slepc_amd import matgen as M

pytestmark = pytest.mark.gpu

EX11 = "0.09789, 0.09789, 0.19577, 0.38197"


@pytest.fixture(scope="module")
def E():
    from slepc_amd import eps as E
    E.initialize(0)
    return E


@pytest.mark.parametrize("pc", ["jacobi", "none"])
def test_ex11_sinvert_golden_vs_oracle(E, oracle, pc):
    csr = M.grid_graph_laplacian(10)
    ones = np.ones((100, 1))
    r = E.solve(csr, nev=4, ncv=20, which=E.EPS.TARGET_MAGNITUDE, target=0.0, deflation=ones,
                st=dict(st_type="sinvert", pc_type=pc))
    o = oracle.eps_solve(csr, nev=4, ncv=20, which=oracle.TARGET_MAGNITUDE, target=0.0, deflation=ones, sinvert=True,
                         pc_none=(pc == "none"))
    print(f"pc={pc} its={r['its']}/{o['its']} ksp_its={r['ksp_its']}/{o['ksp_its']}", flush=True)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == EX11
    assert r["its"] == o["its"]
    assert abs(r["ksp_its"] - o["ksp_its"]) <= 0.02 * o["ksp_its"] + 2
    assert np.allclose(r["eigr"][:4], o["eigr"][:4], rtol=1e-7, atol=0)
    assert (r["err"][:4] < 1e-6).all()


def test_sinvert_closed_form_at_device_size(E):
    # 256 x 255 5-point Laplacian (65,280 rows, simple eigenvalues): the four nearest to a
    # shift below the spectrum, A - sigma I positive definite for CG
    N, Mm, sigma = 256, 255, -0.01
    csr = M.laplacian_2d(N, Mm)
    n = len(csr[0]) - 1
    r = E.solve(csr, nev=4, ncv=20, tol=1e-10, which=E.EPS.TARGET_MAGNITUDE, target=sigma,
                v0=M.uniform_vector(2, n), want_vectors=False, st=dict(st_type="sinvert", ksp_rtol=1e-12))
    closed = np.sort(M.laplacian_2d_eigenvalues(N, Mm))
    near = closed[np.argsort(np.abs(closed - sigma))][:4]
    print(f"n={n} its={r['its']} ksp_its={r['ksp_its']} eig={r['eigr'][:4]}", flush=True)
    assert r["reason"] == 1
    assert np.allclose(r["eigr"][:4], near, rtol=1e-9, atol=0)
    # tol bounds the residual for theta = 1/(lambda - sigma); for A it grows by about
    # |lambda - sigma| |A - sigma I| / |lambda| (|A| <= 8)
    amp = np.abs(near - sigma) * (8 - sigma) / np.abs(near)
    assert (r["err"][:4] < 100 * 1e-10 * amp).all(), (r["err"][:4], amp)   # margin for the close pair


# eps test2 "Interior eigenvalues" (src/eps/tests/test2.c, output/test2_1.out)
TEST2_INTERIOR = {30: "2.10130, 1.89870, 2.30286, 2.50131", 32: "2.09516, 2.28463, 1.90484, 2.47152"}


@pytest.mark.parametrize("n", [30, 32])
def test_interior_indefinite_sinvert_bcgs_vs_oracle(E, oracle, n):
    csr = M.laplacian_1d(n)
    r = E.solve(csr, nev=4, tol=1e-10, which=E.EPS.TARGET_MAGNITUDE, target=2.1,
                st=dict(st_type="sinvert", ksp_type="bcgs", ksp_rtol=1e-12))
    o = oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=2.1, sinvert=True, ksp="bcgs",
                         ksp_rtol=1e-12)
    print(f"n={n} its={r['its']}/{o['its']} ksp_its={r['ksp_its']}/{o['ksp_its']} eig={r['eigr'][:4]}", flush=True)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == TEST2_INTERIOR[n]
    assert np.allclose(r["eigr"][:4], o["eigr"][:4], rtol=1e-10, atol=0)
    closed = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    near = closed[np.argsort(np.abs(closed - 2.1))][:4]
    assert np.allclose(r["eigr"][:4], near, rtol=1e-10, atol=0)
    assert abs(r["its"] - o["its"]) <= 1
    assert (r["err"][:4] < 1e-9).all() and r["ksp_its"] > 0


def test_interior_sinvert_bcgs_2d_vs_oracle(E, oracle):
    # 64 x 63 5-point Laplacian (4,032 rows) with the shift between the 4th and 5th
    # eigenvalues: A - sigma I has four negative eigenvalues, CG cannot run and
    # Jacobi-preconditioned BiCGStab needs about 1,600 iterations per inner solve
    N, Mm = 64, 63
    csr = M.laplacian_2d(N, Mm)
    n = len(csr[0]) - 1
    closed = np.sort(M.laplacian_2d_eigenvalues(N, Mm))
    sigma = 0.5 * (closed[3] + closed[4])
    kw = dict(nev=4, ncv=24, tol=1e-10, target=sigma, v0=M.uniform_vector(3, n), want_vectors=False)
    r = E.solve(csr, which=E.EPS.TARGET_MAGNITUDE, st=dict(st_type="sinvert", ksp_type="bcgs", ksp_rtol=1e-12), **kw)
    o = oracle.eps_solve(csr, which=oracle.TARGET_MAGNITUDE, sinvert=True, ksp="bcgs", ksp_rtol=1e-12, **kw)
    near = closed[np.argsort(np.abs(closed - sigma))][:4]
    print(f"n={n} sigma={sigma:.6f} its={r['its']}/{o['its']} ksp_its={r['ksp_its']}/{o['ksp_its']}", flush=True)
    assert r["reason"] == 1 and abs(r["its"] - o["its"]) <= 1
    assert np.allclose(r["eigr"][:4], o["eigr"][:4], rtol=1e-10, atol=0)
    assert np.allclose(np.sort(r["eigr"][:4]), np.sort(near), rtol=1e-10, atol=0)
    # BiCGStab's iteration count on this ill-conditioned indefinite system follows the
    # rounding of its inner products (the device adds them in another order than the
    # oracle's sequential loop); the outer solve and the eigenvalues do not
    assert abs(r["ksp_its"] - o["ksp_its"]) <= 0.25 * o["ksp_its"]
    with pytest.raises(E.SlepcError):   # CG stops on the indefinite shifted matrix
        E.solve(csr, which=E.EPS.TARGET_MAGNITUDE, st=dict(st_type="sinvert", ksp_type="cg"), **kw)


@pytest.mark.parametrize("n", [30, 32])
def test_interior_bcgs_ilu_vs_oracle(E, oracle, n):
    # ILU(0) factored on the device; on the tridiagonal A - 2.1 I it is the exact LU,
    # so every inner solve takes one BiCGStab iteration
    csr = M.laplacian_1d(n)
    st = dict(st_type="sinvert", ksp_type="bcgs", pc_type="ilu", ksp_rtol=1e-12)
    r = E.solve(csr, nev=4, tol=1e-10, which=E.EPS.TARGET_MAGNITUDE, target=2.1, st=st)
    o = oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=2.1, sinvert=True, ksp="bcgs",
                         pc="ilu", ksp_rtol=1e-12)
    print(f"n={n} its={r['its']}/{o['its']} ksp_its={r['ksp_its']}/{o['ksp_its']}", flush=True)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == TEST2_INTERIOR[n]
    assert np.allclose(r["eigr"][:4], o["eigr"][:4], rtol=1e-10, atol=0)
    assert abs(r["its"] - o["its"]) <= 1 and r["ksp_its"] == o["ksp_its"]


def test_interior_sinvert_bcgs_ilu_2d_vs_oracle(E, oracle):
    N, Mm = 64, 63
    csr = M.laplacian_2d(N, Mm)
    n = len(csr[0]) - 1
    closed = np.sort(M.laplacian_2d_eigenvalues(N, Mm))
    sigma = 0.5 * (closed[3] + closed[4])
    kw = dict(nev=4, ncv=24, tol=1e-10, target=sigma, v0=M.uniform_vector(3, n), want_vectors=False)
    r = E.solve(csr, which=E.EPS.TARGET_MAGNITUDE,
                st=dict(st_type="sinvert", ksp_type="bcgs", pc_type="ilu", ksp_rtol=1e-12), **kw)
    o = oracle.eps_solve(csr, which=oracle.TARGET_MAGNITUDE, sinvert=True, ksp="bcgs", pc="ilu", ksp_rtol=1e-12,
                         **kw)
    near = closed[np.argsort(np.abs(closed - sigma))][:4]
    print(f"n={n} its={r['its']}/{o['its']} ksp_its={r['ksp_its']}/{o['ksp_its']}", flush=True)
    assert r["reason"] == 1 and abs(r["its"] - o["its"]) <= 1
    assert np.allclose(r["eigr"][:4], o["eigr"][:4], rtol=1e-10, atol=0)
    assert np.allclose(np.sort(r["eigr"][:4]), np.sort(near), rtol=1e-10, atol=0)
    assert abs(r["ksp_its"] - o["ksp_its"]) <= 0.25 * o["ksp_its"]


def test_ilu_needs_bcgs(E):
    with pytest.raises(E.SlepcError):
        E.solve(M.laplacian_1d(50), nev=2, which=E.EPS.TARGET_MAGNITUDE, target=-0.1,
                st=dict(st_type="sinvert", ksp_type="cg", pc_type="ilu"))


def test_bcgs_matches_cg_on_a_definite_shift(E):
    csr = M.grid_graph_laplacian(10)
    ones = np.ones((100, 1))
    kw = dict(nev=4, ncv=20, tol=1e-10, which=E.EPS.TARGET_MAGNITUDE, target=-0.05, deflation=ones)
    a = E.solve(csr, st=dict(st_type="sinvert", ksp_type="cg", ksp_rtol=1e-12), **kw)
    b = E.solve(csr, st=dict(st_type="sinvert", ksp_type="bcgs", ksp_rtol=1e-12), **kw)
    assert a["reason"] == 1 and b["reason"] == 1
    assert np.allclose(a["eigr"][:4], b["eigr"][:4], rtol=1e-10, atol=1e-12)


def test_sinvert_requires_target_which(E):
    with pytest.raises(E.SlepcError):
        E.solve(M.laplacian_1d(50), nev=2, which=E.EPS.LARGEST_MAGNITUDE, st=dict(st_type="sinvert"))


def test_sinvert_indefinite_shift_reports_failed_solve(E):
    # A - 2 I of the 1-D Laplacian is indefinite: CG stops with p^T (A - sigma I) p <= 0 (or
    # runs out of iterations) and the solve fails with an error instead of a wrong answer
    with pytest.raises(E.SlepcError):
        E.solve(M.laplacian_1d(200), nev=2, which=E.EPS.TARGET_MAGNITUDE, target=2.0,
                st=dict(st_type="sinvert", ksp_maxits=500))
