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
                st=dict(st_type="sinvert", ksp_type="cg", pc_type=pc))
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
                v0=M.uniform_vector(2, n), want_vectors=False, st=dict(st_type="sinvert", ksp_type="cg", ksp_rtol=1e-12))
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
                st=dict(st_type="sinvert", ksp_type="cg", ksp_maxits=500))


# ---- the direct inner solve, KSPPREONLY + PCLU: the reference's default for STSINVERT
# (stsles.c:47-54), here the dense LU with partial pivoting of ks_lu_create / ks_lu_solve

@pytest.fixture(scope="module")
def ks():
    from slepc_amd import ks as K
    return K


@pytest.mark.parametrize("n,sigma,kind", [(1, 0.5, "rand"), (63, 0.0, "rand"), (64, 0.0, "rand"), (200, 2.0, "lap"),
                                          (1000, 0.3, "rand"), (4096, 2.0, "lap")])
def test_lu_solve_vs_oracle_factors(ks, oracle, n, sigma, kind):
    # ragged sizes (the factors are padded to a multiple of 64), a 1-D Laplacian shifted to a
    # zero diagonal (every column pivots), random nonsymmetric rows; two right-hand sides on
    # the same factors, each against the oracle's factors and the backward error bound
    import scipy.linalg as sl
    csr = M.laplacian_1d(n) if kind == "lap" else M.random_sparse_nonsymmetric(n, min(1.0, 8.0 / n), seed=n)
    A = M.csr_to_dense(*csr) - sigma * np.eye(n)
    lu, ipiv = oracle.lu(csr, sigma)
    ctx = ks.Context(0)
    dA = ks.CSR(ctx, csr)
    F = ks.LU(ctx, dA, sigma)
    rng = np.random.default_rng(n)
    cond = np.linalg.cond(A, 1) if n <= 1000 else 1e4   # the shifted Laplacian of 4,096 rows: about 5e3
    for rhs in range(2):
        b = rng.standard_normal(n)
        db = ks.DeviceArray.from_numpy(ctx, b)
        dx = ks.DeviceArray(ctx, (n,))
        F.solve(db, dx)
        ctx.sync()
        x = dx.numpy()
        xr = sl.lu_solve((lu, ipiv), b)
        bw = np.abs(A @ x - b).max() / (np.abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
        fw = np.abs(x - xr).max() / np.abs(xr).max()
        print(f"n={n} {kind} rhs={rhs} backward={bw:.2e} forward={fw:.2e} cond={cond:.1e}", flush=True)
        assert bw < 1e-14 * max(10, n ** 0.5)
        assert fw < 1e-14 * cond * 100
        db.free()
        dx.free()
    F.check()   # every bounded wait of the pipelined triangular solves was met
    F.close()
    dA.free()
    ctx.close()


def test_lu_cooperative_and_column_step_panels_agree(ks, monkeypatch):
    # the sub-panels factored by one cooperative launch each (the default) and by one launch
    # per column (KS_LU_COOPERATIVE=0) do the same operations in the same order
    n, sigma = 3000, 0.3
    csr = M.random_sparse_nonsymmetric(n, 8.0 / n, seed=7)
    b = M.uniform_vector(5, n)
    ctx = ks.Context(0)
    dA = ks.CSR(ctx, csr)
    db = ks.DeviceArray.from_numpy(ctx, b)
    xs = []
    for coop in ("1", "0"):
        monkeypatch.setenv("KS_LU_COOPERATIVE", coop)
        F = ks.LU(ctx, dA, sigma)
        dx = ks.DeviceArray(ctx, (n,))
        F.solve(db, dx)
        F.check()
        xs.append(dx.numpy())
        dx.free()
        F.close()
    assert np.array_equal(xs[0], xs[1])
    A = M.csr_to_dense(*csr) - sigma * np.eye(n)
    assert np.abs(A @ xs[0] - b).max() / (np.abs(A).sum(axis=1).max() * np.abs(xs[0]).max()) < 1e-13
    db.free()
    dA.free()
    ctx.close()


def test_lu_at_its_largest_size(ks):
    # n = 32,768 (the largest the dense factorization takes; 8.6 GB of factors): the 1-D
    # Laplacian shifted to a zero diagonal, against the banded LAPACK solve
    import scipy.linalg as sl
    n, sigma = 32768, 2.0 + 1e-3
    csr = M.laplacian_1d(n)
    ctx = ks.Context(0)
    dA = ks.CSR(ctx, csr)
    F = ks.LU(ctx, dA, sigma)
    b = M.uniform_vector(11, n)
    db = ks.DeviceArray.from_numpy(ctx, b)
    dx = ks.DeviceArray(ctx, (n,))
    F.solve(db, dx)
    ctx.sync()
    x = dx.numpy()
    ab = np.zeros((3, n))
    ab[0, 1:] = -1.0
    ab[1, :] = 2.0 - sigma
    ab[2, :-1] = -1.0
    xr = sl.solve_banded((1, 1), ab, b)
    r = 2.0 * x - sigma * x
    r[1:] -= x[:-1]
    r[:-1] -= x[1:]
    bw = np.abs(r - b).max() / (4.0 * np.abs(x).max() + np.abs(b).max())
    fw = np.abs(x - xr).max() / np.abs(xr).max()
    print(f"n={n} backward={bw:.2e} forward={fw:.2e}", flush=True)
    assert bw < 1e-14 * n ** 0.5 and fw < 1e-9
    F.check()
    F.close()
    db.free()
    dx.free()
    dA.free()
    ctx.close()


def test_lu_zero_pivot_reported(ks, E):
    # diag(1, ..., 10) - 3 I has a zero column: the factorization reports the zero pivot
    # (KS_ERR_PIVOT) and so does the eigensolve that asked for it
    n = 10
    csr = (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.arange(1.0, n + 1))
    ctx = ks.Context(0)
    dA = ks.CSR(ctx, csr)
    with pytest.raises(ks.KSError):
        ks.LU(ctx, dA, 3.0)
    dA.free()
    ctx.close()
    with pytest.raises(E.SlepcError):
        E.solve(csr, nev=2, which=E.EPS.TARGET_MAGNITUDE, target=3.0, st=dict(st_type="sinvert"))


def test_lu_size_limit_names_the_iterative_solvers(E):
    with pytest.raises(E.SlepcError):
        E.solve(M.laplacian_1d(40000), nev=2, which=E.EPS.TARGET_MAGNITUDE, target=-0.01, st=dict(st_type="sinvert"))


@pytest.mark.parametrize("n", [30, 32])
def test_interior_direct_default_vs_oracle(E, oracle, n):
    # eps test2's interior goldens with the default inner solve
    csr = M.laplacian_1d(n)
    r = E.solve(csr, nev=4, tol=1e-10, which=E.EPS.TARGET_MAGNITUDE, target=2.1, st=dict(st_type="sinvert"))
    o = oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=2.1, sinvert=True,
                         ksp="preonly", pc="lu")
    print(f"n={n} its={r['its']}/{o['its']} solves={r['ksp_its']}/{o['ksp_its']}", flush=True)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == TEST2_INTERIOR[n]
    assert np.allclose(r["eigr"][:4], o["eigr"][:4], rtol=1e-12, atol=0)
    assert r["its"] == o["its"] and r["ksp_its"] == o["ksp_its"]


def test_ex11_direct_vs_oracle(E, oracle):
    # ex11's shift 0 is an eigenvalue (the deflated constant vector): the LU meets a zero
    # pivot there; a shift just below it gives ex11's eigenvalues
    csr = M.grid_graph_laplacian(10)
    ones = np.ones((100, 1))
    kw = dict(nev=4, ncv=20, deflation=ones)
    with pytest.raises(E.SlepcError):
        E.solve(csr, which=E.EPS.TARGET_MAGNITUDE, target=0.0, st=dict(st_type="sinvert"), **kw)
    r = E.solve(csr, which=E.EPS.TARGET_MAGNITUDE, target=-0.01, st=dict(st_type="sinvert"), **kw)
    o = oracle.eps_solve(csr, which=oracle.TARGET_MAGNITUDE, target=-0.01, sinvert=True, ksp="preonly", pc="lu", **kw)
    assert r["reason"] == 1 and ", ".join("%.5f" % v for v in r["eigr"][:4]) == EX11
    print(f"its={r['its']}/{o['its']} err={r['err'][:4]}", flush=True)
    assert r["its"] == o["its"]
    assert np.allclose(r["eigr"][:4], o["eigr"][:4], rtol=1e-10, atol=0)
    # the outer tolerance bounds the residual of theta = 1/(lambda - sigma); for A it grows by
    # about |lambda - sigma| |A - sigma I| / |lambda| (near 10 here)
    assert (r["err"][:4] < 1e-6).all()


def test_markov_nonsymmetric_direct_vs_oracle(E, oracle):
    # eps test11's Markov model (m = 15, N = 120) near 0.5: A - 0.5 I nonsymmetric and
    # indefinite; the four eigenvalues nearest the shift against numpy's
    csr = M.markov_model(15)
    w = np.linalg.eigvals(M.csr_to_dense(*csr))
    near = np.sort(w[np.argsort(np.abs(w - 0.5))][:4].real)
    r = E.solve(csr, nev=4, tol=1e-10, which=E.EPS.TARGET_MAGNITUDE, target=0.5, hermitian=False,
                v0=np.ones(120), st=dict(st_type="sinvert"))
    o = oracle.eps_solve(csr, nev=4, tol=1e-10, which=oracle.TARGET_MAGNITUDE, target=0.5, hermitian=False,
                         v0=np.ones(120), sinvert=True, ksp="preonly", pc="lu")
    print(f"its={r['its']}/{o['its']} eig={r['eigr'][:4]}", flush=True)
    assert r["reason"] == 1 and r["its"] == o["its"]
    assert np.allclose(np.sort(r["eigr"][:4]), near, rtol=1e-10, atol=0)
    assert np.allclose(r["eigr"][:4], o["eigr"][:4], rtol=1e-10, atol=0)


def test_deep_interior_direct_at_device_size(E):
    # 128 x 127 5-point Laplacian (16,256 rows), the shift halfway between its 2,000th and
    # 2,001st eigenvalues: 2,000 negative eigenvalues in A - sigma I, out of reach of the
    # iterative inner solvers; the dense LU (16,256^2, 2.1 GB) solves it exactly
    import time
    N, Mm = 128, 127
    csr = M.laplacian_2d(N, Mm)
    n = len(csr[0]) - 1
    closed = np.sort(M.laplacian_2d_eigenvalues(N, Mm))
    sigma = 0.5 * (closed[1999] + closed[2000])
    t0 = time.time()
    r = E.solve(csr, nev=6, ncv=24, tol=1e-10, which=E.EPS.TARGET_MAGNITUDE, target=sigma,
                v0=M.uniform_vector(5, n), want_vectors=False, st=dict(st_type="sinvert"))
    near = np.sort(closed[np.argsort(np.abs(closed - sigma))][:6])
    print(f"n={n} sigma={sigma:.6f} its={r['its']} solves={r['ksp_its']} {time.time() - t0:.2f}s "
          f"eig={r['eigr'][:6]}", flush=True)
    assert r["reason"] == 1
    assert np.allclose(np.sort(r["eigr"][:6]), near, rtol=1e-10, atol=0)
