"""The seeded workload generators (CPU): sizes, structure and the properties
the workloads rely on."""
import numpy as np

from slepc_amd import matgen as M


def _check_csr(csr, n):
    rowptr, col, val = csr
    assert rowptr.dtype == np.int32 and col.dtype == np.int32 and val.dtype == np.float64
    assert len(rowptr) == n + 1 and rowptr[0] == 0 and rowptr[-1] == len(col) == len(val)
    for i in range(0, n, max(1, n // 97)):
        c = col[rowptr[i]:rowptr[i + 1]]
        assert np.all(np.diff(c) > 0) and c.min() >= 0 and c.max() < n


def test_laplacian_2d_matches_ex2_stencil():
    csr = M.laplacian_2d(100)
    _check_csr(csr, 10000)
    assert csr[0][-1] == 49600   # 5n - 4*100
    A = M.csr_to_dense(*M.laplacian_2d(6))
    assert np.array_equal(A, A.T) and np.all(np.diag(A) == 4)
    assert np.allclose(np.sort(np.linalg.eigvalsh(A))[::-1], M.laplacian_2d_eigenvalues(6))


def test_varcoef_3d_structure():
    N = 12
    csr = M.varcoef_laplacian_3d(N, seed=1)
    n = N ** 3
    _check_csr(csr, n)
    assert csr[0][-1] == 7 * n - 6 * N * N
    A = M.csr_to_dense(*csr)
    assert np.allclose(A, A.T)
    # diagonal = sum of the cell's six face coefficients (boundary faces included) -> SPD
    off = A - np.diag(np.diag(A))
    slack = np.diag(A) + off.sum(axis=1)      # boundary faces only: zero for interior cells
    assert np.all(slack >= -1e-12 * np.diag(A))
    assert np.count_nonzero(slack > 1e-12) == N ** 3 - (N - 2) ** 3
    assert np.linalg.eigvalsh(A).min() > 0
    # deterministic for a seed, different across seeds
    assert np.array_equal(M.varcoef_laplacian_3d(N, seed=1)[2], csr[2])
    assert not np.array_equal(M.varcoef_laplacian_3d(N, seed=2)[2], csr[2])


def test_varcoef_3d_full_size_counts():
    # the 128^3 workload: n = 2,097,152 and nnz = 14,581,760
    N = 128
    n = N ** 3
    assert 7 * n - 6 * N * N == 14581760


def test_convection_diffusion_nonsymmetric_and_normal_enough():
    N = 30
    csr = M.convection_diffusion_2d(N, 10.0, seed=2)
    _check_csr(csr, N * N)
    A = M.csr_to_dense(*csr)
    assert not np.allclose(A, A.T)
    # diffusion-dominated: rows are weakly diagonally dominant with positive diagonal
    off = np.abs(A - np.diag(np.diag(A))).sum(axis=1)
    assert np.all(np.diag(A) >= off - 1e-9 * np.diag(A))
