"""Two-sided Krylov-Schur on the MI355X (EPSSetTwoSided: left eigenvectors from
an Arnoldi process on A^T, ks-twosided.c) and the device transpose it runs on
(ks_csr_transpose + the row-blocked SpMV, through MatMultTranspose): the
reference's ex5 and ex9 goldens with -eps_two_sided, the device against the
oracle (restart counts, eigenvalues, left vectors), left residuals at a
device-sized operator, and A^T x against numpy, bit-for-bit repeatable."""
import numpy as np
import pytest

from slepc_amd import matgen as M

pytestmark = pytest.mark.gpu

EX5 = "1.00000, 0.97137, 0.90423, 0.85714"
EX9 = "0.00007+2.13946i, 0.00007-2.13946i, -0.67386+2.52812i, -0.67386-2.52812i"
EPS = np.finfo(float).eps


@pytest.fixture(scope="module")
def E():
    from slepc_amd import eps as E
    E.initialize(0)
    return E


def _long_row_nonsymmetric(n=5000, seed=8):
    # empty rows and columns, one dense row and one dense column, short rows elsewhere
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if i % 101 == 7:
            cols = np.array([], dtype=int)
        elif i == 33:
            cols = np.sort(rng.choice(n, 3000, replace=False))
        else:
            cols = np.sort(rng.choice(n, rng.integers(1, 9), replace=False))
            if 900 not in cols and i % 2 == 0:
                cols = np.sort(np.append(cols, 900))      # column 900: about n/2 entries
        rows.append(cols)
    rowptr = np.zeros(n + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32)
    val = rng.standard_normal(len(col))
    return rowptr, col, val


@pytest.mark.parametrize("name", ["longrows", "convdiff", "markov"])
def test_mult_transpose_vs_numpy(E, name):
    csr = {"longrows": _long_row_nonsymmetric, "convdiff": lambda: M.convection_diffusion_2d(64, 10.0, seed=3),
           "markov": lambda: M.markov_model(40)}[name]()
    n = len(csr[0]) - 1
    A = E.Mat(csr)
    x = np.random.default_rng(1).standard_normal(n)
    vx, vy = E.Vec.from_array(x), E.Vec(n)
    A.multTranspose(vx, vy)
    y = vy.array()
    rows = np.repeat(np.arange(n), np.diff(csr[0]))
    ref = np.zeros(n)
    np.add.at(ref, csr[1], csr[2] * x[rows])
    bound = np.zeros(n)
    np.add.at(bound, csr[1], np.abs(csr[2] * x[rows]))
    assert np.all(np.abs(y - ref) <= 64 * EPS * bound + 1e-300)
    # repeatable bit for bit: the transpose is deterministic and rebuilt identically
    A2 = E.Mat(csr)
    A2.multTranspose(vx, vy)
    assert np.array_equal(vy.array(), y)
    A.destroy()
    A2.destroy()


def _random_pattern(nr, nc, seed, dense_rows=(), dense_cols=()):
    # rows of varied length (some empty), a few dense rows and columns
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(nr):
        if i in dense_rows:
            c = np.arange(nc)
        else:
            k = int(rng.integers(0, 6)) if i % 7 else 0
            c = np.unique(rng.integers(0, nc, k))
            if dense_cols:
                c = np.union1d(c, [d for d in dense_cols if rng.random() < 0.5])
        rows.append(np.asarray(c, dtype=np.int64))
    rowptr = np.zeros(nr + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32) if rowptr[-1] else np.zeros(0, dtype=np.int32)
    return rowptr, col, rng.standard_normal(len(col))


@pytest.mark.parametrize("nr,nc,kind", [(1, 1, "plain"), (50, 1, "plain"), (3000, 257, "dense"),
                                        (20000, 70000, "plain"), (5000, 1 << 20, "plain"), (1 << 16, 1 << 16, "dense")])
def test_transpose_entry_for_entry(nr, nc, kind):
    # the device transpose (ks_csr_transpose, hand-written radix sort) against scipy's: the
    # same row pointer, column indices in increasing order and values bit for bit
    import scipy.sparse as sp
    from slepc_amd import ks as K
    dense = dict(dense_rows=(0, nr // 2), dense_cols=(0, nc - 1)) if kind == "dense" else {}
    csr = _random_pattern(nr, nc, seed=nr + nc, **dense)
    ctx = K.Context(0)
    A = K.CSR(ctx, csr, ncols=nc)
    T = A.transpose()
    rp, ci, va = T.download()
    ref = sp.csr_matrix((csr[2], csr[1], csr[0]), shape=(nr, nc)).T.tocsr()
    ref.sort_indices()
    print(f"nr={nr} nc={nc} nnz={len(csr[1])}", flush=True)
    assert np.array_equal(rp, ref.indptr) and np.array_equal(ci, ref.indices) and np.array_equal(va, ref.data)
    T.free()
    A.free()
    ctx.close()


@pytest.mark.parametrize("twosided", [False, True])
@pytest.mark.parametrize("lock", [True, False])
def test_ex5_markov_golden(E, twosided, lock):
    csr = M.markov_model(15)
    v0 = np.zeros(120)
    v0[:3] = 1.0
    r = E.solve(csr, nev=4, which=E.EPS.LARGEST_REAL, hermitian=False, v0=v0, twosided=twosided, lock=lock)
    print(f"twosided={twosided} lock={lock} its={r['its']}", flush=True)
    assert r["reason"] == 1
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == EX5


@pytest.mark.parametrize("twosided", [False, True])
def test_ex9_brusselator_golden(E, twosided):
    r = E.solve(M.brusselator(50), nev=4, which=E.EPS.LARGEST_REAL, hermitian=False, twosided=twosided)
    print(f"twosided={twosided} its={r['its']}", flush=True)
    assert r["reason"] == 1
    assert ", ".join("%.5f%+.5fi" % (a, b) for a, b in zip(r["eigr"][:4], r["eigi"][:4])) == EX9


def test_twosided_vs_oracle(E, oracle):
    csr = M.convection_diffusion_2d(40, 10.0, seed=2)
    n = len(csr[0]) - 1
    kw = dict(nev=6, ncv=24, tol=1e-10, hermitian=False, v0=M.uniform_vector(4, n), twosided=True)
    r = E.solve(csr, which=E.EPS.LARGEST_REAL, **kw)
    o = oracle.eps_solve(csr, which=oracle.LARGEST_REAL, **kw)
    print(f"its device={r['its']} oracle={o['its']}", flush=True)
    assert r["reason"] == 1 and abs(r["its"] - o["its"]) <= 1
    za = r["eigr"][:6] + 1j * r["eigi"][:6]
    zb = o["eigr"][:6] + 1j * o["eigi"][:6]
    assert np.allclose(np.sort_complex(za), np.sort_complex(zb), rtol=1e-10, atol=0)
    A = M.csr_to_dense(*csr)
    for i in range(6):
        if r["eigi"][i] != 0.0:
            continue
        y = r["Y"][:, i] / np.linalg.norm(r["Y"][:, i])
        assert np.linalg.norm(A.T @ y - r["eigr"][i] * y) < 1e-8 * abs(r["eigr"][i])
        yo = o["Y"][:, i] / np.linalg.norm(o["Y"][:, i])
        assert abs(abs(y @ yo) - 1.0) < 1e-8


def test_twosided_at_device_size(E):
    # 512 x 512 convection-diffusion (262,144 rows): left and right Arnoldi on the device,
    # every returned left vector checked with the device transpose product
    csr = M.convection_diffusion_2d(512, 10.0, seed=2)
    n = len(csr[0]) - 1
    r = E.solve(csr, nev=4, ncv=32, tol=1e-9, which=E.EPS.LARGEST_REAL, hermitian=False,
                v0=M.uniform_vector(6, n), twosided=True, max_it=300)
    print(f"n={n} its={r['its']} reason={r['reason']} eig={r['eigr'][:4]}", flush=True)
    assert r["reason"] == 1 and r["nconv"] >= 4
    A = E.Mat(csr)
    for i in range(4):
        if r["eigi"][i] != 0.0:
            continue
        y = r["Y"][:, i]
        vy, vz = E.Vec.from_array(y), E.Vec(n)
        A.multTranspose(vy, vz)
        res = np.linalg.norm(vz.array() - r["eigr"][i] * y) / (np.linalg.norm(y) * abs(r["eigr"][i]))
        assert res < 1e-8, res
        assert r["err"][i] <= 1e-8
    A.destroy()


def test_twosided_refused_for_symmetric(E):
    A = E.Mat(M.laplacian_1d(40))
    eps = E.EPS()
    eps.setOperators(A)
    eps.setProblemType(E.EPS.HEP)
    eps.setTwoSided(True)
    with pytest.raises(E.SlepcError):
        eps.solve()
    eps.destroy()
    A.destroy()
