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


def test_er_laplacian_properties():
    csr = M.er_laplacian(2000, 8.0, seed=3)
    _check_csr(csr, 2000)
    A = M.csr_to_dense(*csr)
    assert np.allclose(A, A.T)
    assert np.abs(A.sum(axis=1)).max() < 1e-12          # Laplacian rows sum to zero
    off = A[~np.eye(2000, dtype=bool)]
    w = -off[off != 0]
    assert w.min() > 0 and w.max() <= 1.0               # weights in (0, 1]
    ndeg = np.count_nonzero(A, axis=1) - 1
    assert abs(ndeg.mean() - 8.0) < 0.5                  # average degree 8
    assert np.linalg.eigvalsh(A).min() > -1e-10          # positive semidefinite
    assert np.array_equal(M.er_laplacian(2000, 8.0, seed=3)[1], csr[1])


def test_rmat_shifted_normalized_laplacian():
    csr = M.rmat_shifted_normalized_laplacian(10, 16, seed=5, dtype=np.float64)
    n = 1 << 10
    _check_csr(csr, n)
    A = M.csr_to_dense(*csr)
    assert np.allclose(A, A.T)
    assert np.allclose(np.diag(A), 1.0)
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > -1e-12 and ev.max() < 2 + 1e-12    # 2I - L_norm has its spectrum in [0, 2]
    # the top eigenvalue is 2 once per connected component with an edge: D^1/2 1 is its eigenvector
    deg = np.count_nonzero(A, axis=1) - 1
    u = np.sqrt(deg.astype(float))
    assert np.allclose(A @ u, 2.0 * u)


def test_rmat_average_degree():
    # 8 n drawn edges, symmetrized: the average degree approaches 16 from below as
    # the duplicates of the skewed quadrant choice thin out with the scale (the
    # 2^24 workload: ~15.5, ~268M stored off-diagonal entries)
    prev = 0.0
    for scale in (12, 14):
        rowptr = M.rmat_shifted_normalized_laplacian(scale, 8, seed=5)[0]
        n = 1 << scale
        deg = (int(rowptr[-1]) - n) / n
        assert prev < deg <= 16.0 and deg > 12.5
        prev = deg


def test_rmat_native_generator_matches_numpy_restatement():
    # the OpenMP generator (slepc_amd/gen/rmat.c) and the numpy restatement of the
    # same counter-based steps build the same CSR, bit for bit
    import os
    from slepc_amd import build
    if not os.path.exists(build.GEN_LIB):
        build.build_gen()
    for scale in (8, 11):
        a = M._rmat_laplacian_native(scale, 16, 5, (0.57, 0.19, 0.19, 0.05))
        b = M._rmat_laplacian_numpy(scale, 16, 5, (0.57, 0.19, 0.19, 0.05))
        b = (b[0], b[1], b[2].astype(np.float32))    # rounded once to FP32, as the generator does
        assert a is not None
        for x, y in zip(a, b):
            assert x.dtype == y.dtype and np.array_equal(x, y)


def test_rmat_quadrant_frequencies():
    # level choices follow (a, b, c, d) = (0.57, 0.19, 0.19, 0.05): the top bit of
    # (src, dst) is (down, right) of the first level
    src, dst = M.rmat_edges(12, 16, seed=3)
    top = 1 << 11
    d, r = (src & 1) > 0, (dst & 1) > 0
    freq = [np.mean(~d & ~r), np.mean(~d & r), np.mean(d & ~r), np.mean(d & r)]
    assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.01), freq
    assert src.max() < 2 * top and dst.max() < 2 * top


def test_uniform_vector_is_the_device_stream(oracle):
    # the same counter-based generator as the oracle's and the device's
    for seed in (0, 1, 123456789):
        assert np.array_equal(M.uniform_vector(seed, 4097), oracle.random_uniform(seed, 4097))


def test_serving_stream_parameters():
    reqs = M.serving_stream(2000, qps=50.0, seed=4)
    n = np.array([r["n"] for r in reqs])
    nev = np.array([r["nev"] for r in reqs])
    assert n.min() >= 1000 and n.max() <= 200000
    assert abs(np.log10(n).mean() - (3 + np.log10(2e5)) / 2) < 0.05    # log-uniform
    assert nev.min() == 4 and nev.max() == 32
    assert all(r["ncv"] == 2 * r["nev"] + 16 for r in reqs)
    gaps = np.diff([r["t_arrival"] for r in reqs])
    assert abs(gaps.mean() - 1 / 50.0) < 0.002                          # Poisson at 50 QPS
