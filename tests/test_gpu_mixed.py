"""Mixed-precision Krylov-Schur on the MI355X: an FP32 basis (FP32 SpMV on an
FP32 copy of the operator, FP32 Gram-Schmidt products, FP64 DGKS scalars and
FP64 projected problem) down to 1e-5, whose Ritz vectors seed the FP64
restarts that reach 1e-10.  Checked against the FP64 oracle and dense LAPACK;
the per-pair error against the oracle is asserted, not only the residual."""
import numpy as np
import pytest

from slepc_amd import matgen as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from slepc_amd import eps as E
    E.initialize(0)
    return E


def test_fp32_basis_phase_alone(E, oracle):
    # an FP32-basis solve reaches 1e-5 and its eigenvalues agree with FP64 to FP32 accuracy
    csr = M.varcoef_laplacian_3d(20, seed=1)
    n = len(csr[0]) - 1
    v0 = oracle.random_uniform(0, n)
    r = E.solve(csr, nev=8, ncv=32, tol=1e-5, v0=v0, prec=E.F32)
    ro = oracle.eps_solve(csr, nev=8, ncv=32, tol=1e-10, v0=v0)
    assert r["reason"] == 1 and r["nconv"] >= 8
    assert np.allclose(r["eigr"][:8], ro["eigr"][:8], rtol=1e-5)


def test_mixed_symmetric_reaches_fp64(E, oracle):
    csr = M.varcoef_laplacian_3d(24, seed=1)
    n = len(csr[0]) - 1
    v0 = oracle.random_uniform(0, n)
    r = E.solve(csr, nev=16, ncv=48, tol=1e-10, v0=v0, mixed_tol=1e-5)
    ro = oracle.eps_solve(csr, nev=16, ncv=48, tol=1e-10, v0=v0)
    assert r["reason"] == 1 and r["its32"] > 0
    # per-eigenpair error against the FP64 oracle
    rel = np.abs(r["eigr"][:16] - ro["eigr"][:16]) / np.abs(ro["eigr"][:16])
    assert rel.max() < 1e-10, rel
    assert (r["err"][:16] <= 1e-10).all()
    s = np.linalg.svd(ro["X"][:, :16].T @ r["X"][:, :16], compute_uv=False)
    assert s.min() > 1 - 1e-9


def test_mixed_nonsymmetric_convection_diffusion(E, oracle):
    # the mixed-precision workload's operator family at 64^2 (largest real part)
    csr = M.convection_diffusion_2d(64, 10.0, 2)
    n = len(csr[0]) - 1
    v0 = oracle.random_uniform(0, n)
    r = E.solve(csr, nev=8, ncv=40, tol=1e-10, v0=v0, which=E.EPS.LARGEST_REAL, hermitian=False,
                mixed_tol=1e-5)
    ro = oracle.eps_solve(csr, nev=8, ncv=40, tol=1e-10, v0=v0, which=oracle.LARGEST_REAL, hermitian=False)
    assert r["reason"] == 1
    A = M.csr_to_dense(*csr)
    ev = np.linalg.eigvals(A)
    ev = ev[np.argsort(-ev.real)]
    got = r["eigr"][:8] + 1j * r["eigi"][:8]
    assert np.allclose(np.sort_complex(got), np.sort_complex(ev[:8]), rtol=1e-10)
    assert np.allclose(r["eigr"][:8], ro["eigr"][:8], rtol=1e-10)
    assert (r["err"][:8] <= 1e-10).all()
