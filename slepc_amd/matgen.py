"""Seeded synthetic CSR matrices for the Krylov-Schur workloads.

Every generator returns ``(rowptr, colidx, val)`` as numpy arrays
(int32, int32, float64) in PETSc SeqAIJ layout: rows in order, column
indices strictly increasing inside a row.  Generators are deterministic for a
given seed on any machine (numpy ``default_rng`` / PCG64).

* ``laplacian_1d``      -- the 1-D Laplacian of src/eps/tests/test2.c:40-46.
* ``laplacian_2d``      -- the 2-D 5-point Laplacian of src/eps/tutorials/ex2.c:44-51
                           (row II = i*n + j, stencil [-1,-1,4,-1,-1], Dirichlet).
* ``varcoef_laplacian_3d`` -- 3-D 7-point variable-coefficient Laplacian,
                           lognormal(0,1) face coefficients, diagonal = sum of the
                           cell's six face coefficients (boundary faces included,
                           which makes the matrix SPD).
* ``convection_diffusion_2d`` -- nonsymmetric 5-point upwind convection-diffusion.
* ``er_laplacian``      -- weighted Erdos-Renyi graph Laplacian D - W (serving stream).
* ``rmat_shifted_normalized_laplacian`` -- 2I - (I - D^-1/2 A D^-1/2) of an R-MAT graph.
* ``serving_stream``    -- the request parameters of a Poisson-arrival serving run.
* ``uniform_vector``    -- U(-1,1) from the counter-based splitmix64 generator of the
                           device (the same numbers as VecSetRandomSeeded).
"""
from __future__ import annotations

import numpy as np


def _stencil_to_csr(cols: np.ndarray, vals: np.ndarray, mask: np.ndarray):
    """cols/vals/mask are (n, s) arrays with each row's entries already in
    increasing column order; masked-out entries are dropped."""
    counts = mask.sum(axis=1, dtype=np.int64)
    rowptr = np.zeros(cols.shape[0] + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    if rowptr[-1] > np.iinfo(np.int32).max:
        raise ValueError("nnz exceeds 32-bit index range")
    return (rowptr.astype(np.int32), cols[mask].astype(np.int32),
            np.ascontiguousarray(vals[mask], dtype=np.float64))


def laplacian_1d(n: int):
    i = np.arange(n, dtype=np.int64)
    cols = np.stack([i - 1, i, i + 1], axis=1)
    vals = np.broadcast_to(np.array([-1.0, 2.0, -1.0]), cols.shape).copy()
    mask = (cols >= 0) & (cols < n)
    return _stencil_to_csr(cols, vals, mask)


def laplacian_2d(n: int, m: int | None = None):
    """ex2.c ordering: II = i*n + j, i in [0,m), j in [0,n)."""
    m = n if m is None else m
    N = n * m
    II = np.arange(N, dtype=np.int64)
    i = II // n
    j = II - i * n
    cols = np.stack([II - n, II - 1, II, II + 1, II + n], axis=1)
    mask = np.stack([i > 0, j > 0, np.ones(N, bool), j < n - 1, i < m - 1], axis=1)
    vals = np.broadcast_to(np.array([-1.0, -1.0, 4.0, -1.0, -1.0]), cols.shape).copy()
    return _stencil_to_csr(cols, vals, mask)


def laplacian_2d_eigenvalues(n: int, m: int | None = None) -> np.ndarray:
    """All eigenvalues of laplacian_2d (closed form), descending."""
    m = n if m is None else m
    p = np.arange(1, n + 1)
    q = np.arange(1, m + 1)
    lp = 4.0 * np.sin(p * np.pi / (2 * (n + 1))) ** 2
    lq = 4.0 * np.sin(q * np.pi / (2 * (m + 1))) ** 2
    return np.sort((lp[:, None] + lq[None, :]).ravel())[::-1]


def varcoef_laplacian_3d(N: int, seed: int = 1):
    """Cell (x,y,z) has row index x + N*(y + N*z).  Face coefficients are
    drawn in the order ax (z,y,x-face), ay (z,y-face,x), az (z-face,y,x),
    each lognormal(0,1) from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    ax = rng.lognormal(0.0, 1.0, size=(N, N, N + 1))
    ay = rng.lognormal(0.0, 1.0, size=(N, N + 1, N))
    az = rng.lognormal(0.0, 1.0, size=(N + 1, N, N))
    diag = (ax[:, :, :-1] + ax[:, :, 1:] + ay[:, :-1, :] + ay[:, 1:, :]
            + az[:-1, :, :] + az[1:, :, :]).ravel()
    n = N ** 3
    idx = np.arange(n, dtype=np.int64)
    x = idx % N
    y = (idx // N) % N
    z = idx // (N * N)
    NN = N * N
    cols = np.stack([idx - NN, idx - N, idx - 1, idx, idx + 1, idx + N, idx + NN], axis=1)
    vals = np.empty((n, 7), dtype=np.float64)
    vals[:, 0] = -az[:-1, :, :].ravel()       # face between z-1 and z
    vals[:, 1] = -ay[:, :-1, :].ravel()
    vals[:, 2] = -ax[:, :, :-1].ravel()
    vals[:, 3] = diag
    vals[:, 4] = -ax[:, :, 1:].ravel()
    vals[:, 5] = -ay[:, 1:, :].ravel()
    vals[:, 6] = -az[1:, :, :].ravel()
    mask = np.stack([z > 0, y > 0, x > 0, np.ones(n, bool), x < N - 1, y < N - 1, z < N - 1], axis=1)
    del ax, ay, az, diag
    return _stencil_to_csr(cols, vals, mask)


def convection_diffusion_2d(N: int, peclet: float = 10.0, seed: int = 2):
    """-div(k grad u) + w . grad u on the unit square, N x N interior cells,
    Dirichlet boundary, first-order upwind convection with constant velocity
    along the diagonal, |w| chosen so that the (global) Peclet number
    |w| L / k_mean equals ``peclet`` (L = 1, k_mean = 1.25); k is uniform(0.5, 2)
    per cell (default_rng(seed)); face diffusion is the harmonic mean of the two
    cells.  Row x + N*y.  (A grid Peclet number of 10 instead would make the
    operator so non-normal that Krylov Ritz values land in the pseudospectrum,
    far from the eigenvalues, at any working precision.)"""
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.5, 2.0, size=(N, N))          # [y, x]
    h = 1.0 / (N + 1)
    kmean = 1.25
    wmag = peclet * kmean
    wx = wy = wmag / np.sqrt(2.0)
    # face diffusion coefficients (harmonic mean; boundary faces use the cell value)
    kx = np.empty((N, N + 1))
    kx[:, 1:-1] = 2.0 * k[:, :-1] * k[:, 1:] / (k[:, :-1] + k[:, 1:])
    kx[:, 0] = k[:, 0]
    kx[:, -1] = k[:, -1]
    ky = np.empty((N + 1, N))
    ky[1:-1, :] = 2.0 * k[:-1, :] * k[1:, :] / (k[:-1, :] + k[1:, :])
    ky[0, :] = k[0, :]
    ky[-1, :] = k[-1, :]
    inv_h2 = 1.0 / (h * h)
    west = -(kx[:, :-1] * inv_h2) - wx / h      # upwind for wx > 0 takes u_{x-1}
    east = -(kx[:, 1:] * inv_h2)
    south = -(ky[:-1, :] * inv_h2) - wy / h
    north = -(ky[1:, :] * inv_h2)
    diag = (kx[:, :-1] + kx[:, 1:] + ky[:-1, :] + ky[1:, :]) * inv_h2 + (wx + wy) / h
    n = N * N
    idx = np.arange(n, dtype=np.int64)
    x = idx % N
    y = idx // N
    cols = np.stack([idx - N, idx - 1, idx, idx + 1, idx + N], axis=1)
    vals = np.stack([south.ravel(), west.ravel(), diag.ravel(), east.ravel(), north.ravel()], axis=1)
    mask = np.stack([y > 0, x > 0, np.ones(n, bool), x < N - 1, y < N - 1], axis=1)
    return _stencil_to_csr(cols, vals, mask)


def csr_to_dense(rowptr, colidx, val, n: int | None = None) -> np.ndarray:
    n = len(rowptr) - 1 if n is None else n
    A = np.zeros((len(rowptr) - 1, n))
    rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
    A[rows, colidx] = val
    return A


def random_sparse_symmetric(n: int, density: float, seed: int):
    """Small symmetric sparse test matrix with a dominant diagonal."""
    rng = np.random.default_rng(seed)
    M = (rng.random((n, n)) < density) * rng.standard_normal((n, n))
    M = np.triu(M, 1)
    M = M + M.T + np.diag(rng.uniform(-3, 3, n))
    return dense_to_csr(M)


def markov_model(m: int):
    """Random walk on a triangular grid with m points a side (N = m(m+1)/2), the
    transpose of a stochastic matrix: the test matrix of the reference's
    src/eps/tutorials/ex5.c:139-170 (MatMarkovModel), built row by row the same way."""
    cst = 0.5 / (m - 1)
    N = m * (m + 1) // 2
    rows = [dict() for _ in range(N)]
    ix = 0
    for i in range(1, m + 1):
        jmax = m - i + 1
        for j in range(1, jmax + 1):
            ix += 1
            r = rows[ix - 1]
            if j != jmax:
                pd = cst * (i + j - 1)
                r[ix] = 2 * pd if i == 1 else pd                     # north
                r[ix + jmax - 1] = 2 * pd if j == 1 else pd          # east
            pu = 0.5 - cst * (i + j - 3)
            if j > 1:
                r[ix - 2] = pu                                       # south
            if i > 1:
                r[ix - jmax - 2] = pu                                # west
    return _rows_to_csr(rows)


def brusselator(n: int, alpha: float = 2.0, beta: float = 5.45, delta1: float = 0.008, delta2: float = 0.004,
                L: float = 0.51302):
    """The Brusselator wave model of src/eps/tutorials/ex9.c:20-35, 92-121 as an
    explicit 2n x 2n matrix [tau1 T + (beta-1) I, alpha^2 I; -beta I, tau2 T - alpha^2 I]
    with T = tridiag(1, -2, 1), h = 1/(n+1), tau_i = delta_i / (h L)^2."""
    h = 1.0 / (n + 1)
    tau1 = delta1 / ((h * L) ** 2)
    tau2 = delta2 / ((h * L) ** 2)
    rows = [dict() for _ in range(2 * n)]
    for i in range(n):
        r1, r2 = rows[i], rows[n + i]
        r1[i] = -2.0 * tau1 + beta - 1.0
        r2[n + i] = -2.0 * tau2 - alpha * alpha
        if i > 0:
            r1[i - 1] = tau1
            r2[n + i - 1] = tau2
        if i < n - 1:
            r1[i + 1] = tau1
            r2[n + i + 1] = tau2
        r1[n + i] = alpha * alpha
        r2[i] = -beta
    return _rows_to_csr(rows)


def _rows_to_csr(rows):
    rowptr = np.zeros(len(rows) + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in sorted(r)], dtype=np.int32)
    val = np.array([r[c] for r in rows for c in sorted(r)], dtype=np.float64)
    return rowptr, col, val


def grid_graph_laplacian(n: int, m: int | None = None):
    """Laplacian L(G) of the n x m grid graph: the degree on the diagonal, -1 per
    edge (src/eps/tutorials/ex11.c:40-60); singular, the constant vector spans
    its null space."""
    m = n if m is None else m
    rows = [dict() for _ in range(n * m)]
    for II in range(n * m):
        i, j = divmod(II, n)
        w = 0.0
        for cond, c in ((i > 0, II - n), (i < m - 1, II + n), (j > 0, II - 1), (j < n - 1, II + 1)):
            if cond:
                rows[II][c] = -1.0
                w += 1.0
        rows[II][II] = w
    return _rows_to_csr(rows)


def random_sparse_nonsymmetric(n: int, density: float, seed: int):
    rng = np.random.default_rng(seed)
    M = (rng.random((n, n)) < density) * rng.standard_normal((n, n))
    M = M + np.diag(rng.uniform(-3, 3, n))
    return dense_to_csr(M)


def dense_to_csr(M: np.ndarray):
    mask = M != 0
    n = M.shape[0]
    counts = mask.sum(axis=1)
    rowptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(counts, out=rowptr[1:])
    cols = np.nonzero(mask)[1].astype(np.int32)
    return rowptr, cols, np.ascontiguousarray(M[mask], dtype=np.float64)


def _csr_from_coo(n: int, r: np.ndarray, c: np.ndarray, v: np.ndarray):
    """CSR with sorted, unique columns per row from COO triplets (no duplicates)."""
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    counts = np.bincount(r, minlength=n).astype(np.int64)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    if rowptr[-1] > np.iinfo(np.int32).max:
        raise ValueError("nnz exceeds 32-bit index range")
    return rowptr.astype(np.int32), c.astype(np.int32), np.ascontiguousarray(v, dtype=np.float64)


def er_laplacian(n: int, avg_degree: float = 8.0, seed: int = 0):
    """Laplacian D - W of a weighted Erdos-Renyi graph G(n, p), p = avg_degree/(n-1),
    edge weights uniform(0, 1].  The edge count is drawn Binomial(n(n-1)/2, p) and
    that many distinct unordered pairs are drawn uniformly (which is G(n, p) in
    distribution).  Every row stores its diagonal, also for isolated vertices."""
    rng = np.random.default_rng(seed)
    npairs = n * (n - 1) // 2
    p = min(1.0, avg_degree / max(n - 1, 1))
    m = int(rng.binomial(npairs, p))
    keys = np.empty(0, dtype=np.int64)
    while len(keys) < m:
        need = m - len(keys)
        a = rng.integers(0, n, size=int(need * 1.1) + 16, dtype=np.int64)
        b = rng.integers(0, n, size=len(a), dtype=np.int64)
        ok = a != b
        lo, hi = np.minimum(a[ok], b[ok]), np.maximum(a[ok], b[ok])
        keys = np.unique(np.concatenate([keys, lo * n + hi]))
    keys = keys[rng.permutation(len(keys))[:m]]
    i, j = keys // n, keys % n
    w = 1.0 - rng.random(m)                      # uniform(0, 1]
    deg = np.bincount(i, weights=w, minlength=n) + np.bincount(j, weights=w, minlength=n)
    r = np.concatenate([i, j, np.arange(n, dtype=np.int64)])
    c = np.concatenate([j, i, np.arange(n, dtype=np.int64)])
    v = np.concatenate([-w, -w, deg])
    return _csr_from_coo(n, r, c, v)


def _rmat_thresholds(abcd):
    a, b, c, _ = abcd
    return (np.uint64(int(round(a * 65536.0))), np.uint64(int(round((a + b) * 65536.0))),
            np.uint64(int(round((a + b + c) * 65536.0))))


def rmat_edges(scale: int, edge_factor: int = 16, abcd=(0.57, 0.19, 0.19, 0.05), seed: int = 5,
               chunk: int = 1 << 20):
    """Directed R-MAT edge list (Graph500 recursive quadrant choice, no noise):
    2^scale vertices, edge_factor * 2^scale edges, as int64 (src, dst) arrays.
    Counter-based randomness: edge e draws r_g = splitmix64(seed * golden + e*G + g),
    g < G = ceil(scale/4); level l = 4g + t uses the 16-bit fraction
    (r_g >> 16t) & 0xffff against round(65536 a), round(65536 (a+b)),
    round(65536 (a+b+c)) -- the same stream as slepc_amd/gen/rmat.c."""
    TA, TB, TC = _rmat_thresholds(abcd)
    G = (scale + 3) // 4
    m = edge_factor << scale
    src = np.empty(m, dtype=np.int64)
    dst = np.empty(m, dtype=np.int64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * _SM_GOLDEN
        for s0 in range(0, m, chunk):
            e1 = min(m, s0 + chunk)
            e = np.arange(s0, e1, dtype=np.uint64)
            u = np.zeros(e1 - s0, dtype=np.uint64)
            v = np.zeros(e1 - s0, dtype=np.uint64)
            for g in range(G):
                r = _splitmix64(base + e * np.uint64(G) + np.uint64(g))
                for t in range(4):
                    lvl = 4 * g + t
                    if lvl >= scale:
                        break
                    x = (r >> np.uint64(16 * t)) & np.uint64(0xFFFF)
                    right = ((x >= TA) & (x < TB)) | (x >= TC)
                    down = x >= TB
                    u |= down.astype(np.uint64) << np.uint64(lvl)
                    v |= right.astype(np.uint64) << np.uint64(lvl)
            src[s0:e1], dst[s0:e1] = u.astype(np.int64), v.astype(np.int64)
    return src, dst


def _rmat_laplacian_numpy(scale, edge_factor, seed, abcd):
    n = 1 << scale
    src, dst = rmat_edges(scale, edge_factor, abcd, seed)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    del src, dst, keep
    key = np.unique(lo * n + hi)
    del lo, hi
    i, j = key // n, key % n
    del key
    deg = (np.bincount(i, minlength=n) + np.bincount(j, minlength=n)).astype(np.float64)
    dinv = np.where(deg > 0, 1.0 / np.sqrt(np.maximum(deg, 1.0)), 0.0)
    w = dinv[i] * dinv[j]
    r = np.concatenate([i, j, np.arange(n, dtype=np.int64)])
    c = np.concatenate([j, i, np.arange(n, dtype=np.int64)])
    del i, j
    v = np.concatenate([w, w, np.ones(n)])
    return _csr_from_coo(n, r, c, v)   # FP64 values


def _rmat_laplacian_native(scale, edge_factor, seed, abcd):
    """The same matrix from slepc_amd/libks_gen.so (OpenMP); None when it is not built."""
    import ctypes as C
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libks_gen.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ks_gen_rmat_laplacian.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_int64)]
    L.ks_gen_free.argtypes = [C.c_void_p]
    rp, ci, va, nnz = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_int64()
    a, b, c, _ = abcd
    e = L.ks_gen_rmat_laplacian(scale, edge_factor, C.c_uint64(seed), a, b, c, C.byref(rp), C.byref(ci),
                                C.byref(va), C.byref(nnz))
    if e != 0:
        raise MemoryError(f"ks_gen_rmat_laplacian failed ({e})")
    n = 1 << scale
    try:
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int32)), shape=(n + 1,)).copy()
        colidx = np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), shape=(nnz.value,)).copy()
        val = np.ctypeslib.as_array(C.cast(va, C.POINTER(C.c_float)), shape=(nnz.value,)).copy()
    finally:
        for p in (rp, ci, va):
            L.ks_gen_free(p)
    return rowptr, colidx, val


def rmat_shifted_normalized_laplacian(scale: int, edge_factor: int = 16, seed: int = 5,
                                      abcd=(0.57, 0.19, 0.19, 0.05), dtype=np.float32, native: bool = True):
    """M = 2I - (I - D^-1/2 A D^-1/2) = I + D^-1/2 A D^-1/2 of the R-MAT graph
    symmetrized (A = pattern of E + E^T, E the edge_factor * 2^scale drawn
    edges, so the average degree is at most 2 edge_factor), with self-loops
    and duplicate edges dropped; isolated vertices get M_ii = 1 (normalized Laplacian L_ii = 1 for
    them, the convention with D^-1/2 = 0 on isolated vertices).  Off-diagonal
    values are d_i^-1/2 d_j^-1/2 formed in FP64 and rounded once to FP32.
    Eigenvalues lie in [0, 2]; the largest of M are the smallest of the
    normalized Laplacian.  ``native``: an FP32 matrix comes from the OpenMP
    generator in libks_gen.so when that is built (the same matrix, much faster
    at scale); FP64 values are formed in numpy and not rounded."""
    out = None
    if native and np.dtype(dtype) == np.float32:
        out = _rmat_laplacian_native(scale, edge_factor, seed, abcd)
    if out is None:
        out = _rmat_laplacian_numpy(scale, edge_factor, seed, abcd)
    rowptr, colidx, val = out
    return rowptr, colidx, val.astype(dtype)


_SM_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + _SM_GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform_vector(seed: int, n: int) -> np.ndarray:
    """u_i = 2 * (splitmix64(seed * golden + i) >> 11) * 2^-53 - 1: the device's
    counter-based U(-1,1) stream (VecSetRandomSeeded / ks_bv_random)."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * _SM_GOLDEN
        x = base + np.arange(n, dtype=np.uint64)
    return 2.0 * (_splitmix64(x) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 - 1.0


def serving_stream(nreq: int, qps: float, seed: int = 4, nmin: int = 1000, nmax: int = 200000):
    """Request parameters of the serving workload: n = floor(10^U(log10 nmin, log10 nmax)),
    nev uniform in {4..32}, ncv = 2 nev + 16, Poisson arrivals at ``qps`` (exponential
    gaps), a graph seed and a start-vector seed per request."""
    rng = np.random.default_rng(seed)
    n = np.floor(10.0 ** rng.uniform(np.log10(nmin), np.log10(nmax), nreq)).astype(np.int64)
    nev = rng.integers(4, 33, nreq)
    t = np.cumsum(rng.exponential(1.0 / qps, nreq)) if qps > 0 else np.zeros(nreq)
    return [dict(n=int(n[i]), nev=int(nev[i]), ncv=int(2 * nev[i] + 16), t_arrival=float(t[i]),
                 graph_seed=1000 + i, v0_seed=2000 + i) for i in range(nreq)]
