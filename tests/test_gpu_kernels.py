"""Kernel-level parity of libks_hip.so against the CPU oracle and numpy,
called through the C ABI (include/ks_hip.h).  Needs an MI355X."""
import numpy as np
import pytest

from slepc_amd import matgen as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ks():
    from slepc_amd import ks as K
    return K


@pytest.fixture(scope="module")
def ctx(ks):
    c = ks.Context(0)
    yield c
    c.close()


def _rel_bound(csr, x):
    """|A| |x| row-wise: the scale of the rounding error of a row sum."""
    rowptr, col, val = csr
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    out = np.zeros(n)
    np.add.at(out, rows, np.abs(val) * np.abs(x[col]))
    return out


def _long_row_matrix(n=3000, seed=4):
    # rows of very different lengths: empty rows, a 2500-entry row, short rows
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if i % 97 == 5:
            cols = np.array([], dtype=int)
        elif i == 17:
            cols = np.sort(rng.choice(n, 2500, replace=False))
        elif i % 50 == 0:
            cols = np.sort(rng.choice(n, 300, replace=False))
        else:
            cols = np.sort(rng.choice(n, rng.integers(1, 12), replace=False))
        rows.append(cols)
    rowptr = np.zeros(n + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32)
    val = rng.standard_normal(len(col))
    return rowptr, col, val


@pytest.mark.parametrize("name", ["lap2d", "varcoef3d", "longrows", "nonsym"])
def test_spmv_f64_vs_oracle(ks, ctx, oracle, name):
    csr = {"lap2d": lambda: M.laplacian_2d(100), "varcoef3d": lambda: M.varcoef_laplacian_3d(20, seed=1),
           "longrows": _long_row_matrix, "nonsym": lambda: M.random_sparse_nonsymmetric(2000, 0.01, 9)}[name]()
    n = len(csr[0]) - 1
    A = ks.CSR(ctx, csr)
    info = A.info()
    if name == "longrows":
        assert info["max_row"] == 2500 and info["variant"] == 1
    x = np.random.default_rng(2).standard_normal(n)
    dx = ks.DeviceArray.from_numpy(ctx, x)
    dy = ks.DeviceArray(ctx, (n,))
    ks.spmv(ctx, A, dx, dy)
    y = dy.numpy()
    yref = oracle.spmv(csr, x)
    bound = _rel_bound(csr, x)
    # summation order differs from the CPU row loop: error <= ~nnz_row * eps * |A||x|
    assert np.all(np.abs(y - yref) <= 64 * np.finfo(float).eps * bound + 1e-300)
    A.free()


def test_spmv_closed_form_sine_mode(ks, ctx):
    n, p, q = 100, 3, 7
    i = np.arange(n)
    u = np.outer(np.sin((i + 1) * q * np.pi / (n + 1)), np.sin((i + 1) * p * np.pi / (n + 1))).ravel()
    lam = 4 * np.sin(p * np.pi / (2 * (n + 1))) ** 2 + 4 * np.sin(q * np.pi / (2 * (n + 1))) ** 2
    A = ks.CSR(ctx, M.laplacian_2d(n))
    du = ks.DeviceArray.from_numpy(ctx, u)
    dy = ks.DeviceArray(ctx, (n * n,))
    ks.spmv(ctx, A, du, dy)
    assert np.abs(dy.numpy() - lam * u).max() < 1e-13


def test_spmv_f32(ks, ctx):
    csr = M.varcoef_laplacian_3d(16, seed=3)
    n = len(csr[0]) - 1
    A = ks.CSR(ctx, csr, prec=ks.F32)
    x = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    dx = ks.DeviceArray.from_numpy(ctx, x)
    dy = ks.DeviceArray(ctx, (n,), np.float32)
    ks.spmv(ctx, A, dx, dy, prec=ks.F32)
    yref = M.csr_to_dense(csr[0], csr[1], csr[2].astype(np.float32).astype(np.float64)) @ x.astype(np.float64)
    bound = _rel_bound((csr[0], csr[1], csr[2]), x.astype(np.float64))
    assert np.all(np.abs(dy.numpy() - yref) <= 16 * np.finfo(np.float32).eps * bound + 1e-30)


@pytest.mark.parametrize("n,k", [(1000, 1), (10000, 20), (70001, 48), (300000, 129)])
def test_dotvec_multvec(ks, ctx, n, k):
    rng = np.random.default_rng(n + k)
    ld = ((n + 63) // 64) * 64
    V = np.asfortranarray(rng.standard_normal((ld, k + 1)))
    dV = ks.DeviceArray.from_numpy(ctx, V)
    y = V[:n, k].copy()
    dq = ks.DeviceArray(ctx, (k,))
    ks.dotvec(ctx, ks.F64, n, dV, ld, k, dV, dq, yoff=k * ld)
    q = dq.numpy()
    ref = V[:n, :k].T @ y
    scale = np.abs(V[:n, :k]).T @ np.abs(y)
    assert np.all(np.abs(q - ref) <= 8 * n ** 0.5 * np.finfo(float).eps * scale)
    # y = 0.5 y - 2 V q
    ks.multvec(ctx, ks.F64, n, -2.0, dV, ld, k, dq, 0.5, dV, yoff=k * ld)
    got = dV.numpy()[:n, k]
    want = 0.5 * y - 2.0 * V[:n, :k] @ q
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("n,K,s,e", [(1000, 20, 3, 12), (4097, 48, 0, 24), (64, 4, 0, 4),
                                     (100003, 97, 5, 101), (20000, 129, 0, 129)])
def test_multinplace_mfma_vs_numpy(ks, ctx, n, K, s, e):
    # V(:, s:e) = V(:, 0:K) Q(0:K, s:e), in place (output columns overlap the input)
    rng = np.random.default_rng(n)
    ld = ((n + 63) // 64) * 64
    ncol = max(K, e)
    V = np.asfortranarray(rng.standard_normal((ld, ncol)))
    Q = np.asfortranarray(rng.standard_normal((ncol + 1, ncol + 1)))
    dV = ks.DeviceArray.from_numpy(ctx, V)
    dQ = ks.DeviceArray.from_numpy(ctx, Q)
    ks.multinplace(ctx, ks.F64, n, dV, ld, K, dQ, ncol + 1, s, e)
    got = dV.numpy()
    want = V.copy()
    want[:n, s:e] = V[:n, :K] @ Q[:K, s:e]
    scale = np.abs(V[:n, :K]) @ np.abs(Q[:K, s:e])
    assert np.all(np.abs(got[:n, s:e] - want[:n, s:e]) <= 4 * K * np.finfo(float).eps * scale + 1e-300)
    keep = [c for c in range(ncol) if not (s <= c < e)]
    assert np.array_equal(got[:n, keep], V[:n, keep])
    assert np.array_equal(got[n:, :], V[n:, :])   # padding rows untouched


def test_multinplace_identity_exact(ks, ctx):
    # Q = permutation: the MFMA path must reproduce the columns bit for bit
    n, K = 5000, 32
    rng = np.random.default_rng(1)
    V = np.asfortranarray(rng.standard_normal((n, K)))
    P = np.eye(K)[:, rng.permutation(K)]
    dV = ks.DeviceArray.from_numpy(ctx, V)
    dQ = ks.DeviceArray.from_numpy(ctx, np.asfortranarray(P))
    ks.multinplace(ctx, ks.F64, n, dV, n, K, dQ, K, 0, K)
    assert np.array_equal(dV.numpy(), V @ P)


def _orth_case(ks, ctx, oracle, V, j, refine=0):
    n = V.shape[0]
    Vo = np.asfortranarray(V.copy())
    H, nrm, lindep, passes = oracle.orthonormalize_column(Vo, j, refine=refine)
    dV = ks.DeviceArray.from_numpy(ctx, V)
    ldbuf = j + 2
    buf = ks.DeviceArray(ctx, (ldbuf, j + 2))
    ctx.status_reset()
    ks.orthonormalize_column(ctx, ks.F64, n, dV, n, j, buf, ldbuf, j, j, refine=refine)
    st = ctx.status()
    Vg = dV.numpy()
    h = buf.numpy()[:, j]
    return dict(H=H, nrm=nrm, lindep=lindep, passes=passes, Vo=Vo, Vg=Vg, h=h, st=st)


def test_orthonormalize_vs_oracle(ks, ctx, oracle):
    rng = np.random.default_rng(11)
    n, j = 50000, 30
    V = np.zeros((n, j + 1), order="F")
    V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
    V[:, j] = rng.standard_normal(n)
    r = _orth_case(ks, ctx, oracle, V, j)
    st = r["st"]
    assert st.total_passes == r["passes"] and st.lindep == int(r["lindep"]) and st.done == 1
    assert abs(st.nrm - r["nrm"]) <= 1e-13 * r["nrm"]
    assert np.allclose(r["h"][:j + 1], r["H"], rtol=1e-10, atol=1e-13)
    assert np.abs(r["Vg"][:, j] - r["Vo"][:, j]).max() < 1e-12
    assert abs(np.linalg.norm(r["Vg"][:, j]) - 1) < 1e-13
    assert np.abs(r["Vg"][:, :j].T @ r["Vg"][:, j]).max() < 1e-14


@pytest.mark.parametrize("j", [0, 1, 9, 16, 17, 24, 48, 96, 128, 129, 200])
def test_orthonormalize_widths(ks, ctx, oracle, j):
    # both Gram-Schmidt kernels: register-resident rows (<= 48 columns) and LDS tiles (more)
    rng = np.random.default_rng(100 + j)
    n = 30011
    V = np.zeros((n, j + 1), order="F")
    if j:
        V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
    V[:, j] = rng.standard_normal(n) + (V[:, :j] @ rng.standard_normal(j) if j else 0.0)
    r = _orth_case(ks, ctx, oracle, V, j)
    st = r["st"]
    assert st.total_passes == r["passes"] and st.lindep == int(r["lindep"]) and st.done == 1
    assert abs(st.nrm - r["nrm"]) <= 1e-12 * r["nrm"]
    assert np.allclose(r["h"][:j + 1], r["H"], rtol=1e-10, atol=1e-12)
    assert np.abs(r["Vg"][:, j] - r["Vo"][:, j]).max() < 1e-12
    if j:
        assert np.abs(r["Vg"][:, :j].T @ r["Vg"][:, j]).max() < 1e-14


@pytest.mark.parametrize("prec,j", [("f64", 9), ("f64", 24), ("f64", 48), ("f64", 96), ("f64", 200),
                                    ("f32", 40), ("f32", 150)])
def test_finish_published_equals_ticket_bitwise(ks, ctx, prec, j):
    # the finishing launch publishes the column sums as epoch-tagged words and
    # takes the DGKS step in one more workgroup; a context made with
    # KS_GS_FIN_TICKET=1 counts with a ticket and fences instead: the same sums,
    # so the same column, coefficients and status bit for bit (three calls in a
    # row, so later calls run on advanced epochs)
    import os
    rng = np.random.default_rng(300 + j)
    n = 40009
    dt = np.float64 if prec == "f64" else np.float32
    p = ks.F64 if prec == "f64" else ks.F32
    os.environ["KS_GS_FIN_TICKET"] = "1"
    try:
        ctx_t = ks.Context(0)
    finally:
        del os.environ["KS_GS_FIN_TICKET"]
    try:
        for rep in range(3):
            V = np.zeros((n, j + 1), order="F")
            V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
            V[:, j] = rng.standard_normal(n) + V[:, :j] @ rng.standard_normal(j)
            V = np.asfortranarray(V.astype(dt))
            out = []
            for c in (ctx, ctx_t):
                dV = ks.DeviceArray.from_numpy(c, V)
                buf = ks.DeviceArray(c, (j + 2, j + 2))
                c.status_reset()
                ks.orthonormalize_column(c, p, n, dV, n, j, buf, j + 2, j, j)
                st = c.status()
                out.append((dV.numpy()[:, j], buf.numpy()[:j + 1, j], st.total_passes, st.nrm, st.lindep, st.error))
            (v0, h0, p0, n0, l0, e0), (v1, h1, p1, n1, l1, e1) = out
            assert e0 == 0 and e1 == 0
            assert np.array_equal(v0, v1) and np.array_equal(h0, h1)
            assert (p0, n0, l0) == (p1, n1, l1)
            assert abs(np.linalg.norm(v0.astype(np.float64)) - 1) < (1e-13 if prec == "f64" else 1e-5)
    finally:
        ctx_t.close()


@pytest.mark.parametrize("j", [5, 40, 60, 150])
def test_orthonormalize_fp32(ks, ctx, oracle, j):
    # FP32 basis: FP32 loads/stores, FP64 sums and DGKS scalars; checked against the
    # FP64 oracle run on the same (FP32-representable) data
    rng = np.random.default_rng(200 + j)
    n = 20011
    V = np.zeros((n, j + 1), order="F")
    V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
    V[:, j] = rng.standard_normal(n)
    V = V.astype(np.float32).astype(np.float64)
    Vo = np.asfortranarray(V.copy())
    H, nrm, lindep, passes = oracle.orthonormalize_column(Vo, j)
    dV = ks.DeviceArray.from_numpy(ctx, np.asfortranarray(V.astype(np.float32)))
    buf = ks.DeviceArray(ctx, (j + 2, j + 2))
    ctx.status_reset()
    ks.orthonormalize_column(ctx, ks.F32, n, dV, n, j, buf, j + 2, j, j)
    st = ctx.status()
    Vg = dV.numpy().astype(np.float64)
    h = buf.numpy()[:, j]
    assert st.done == 1 and not st.lindep
    assert abs(st.nrm - nrm) <= 1e-5 * nrm
    assert np.allclose(h[:j + 1], H, rtol=1e-4, atol=1e-5 * nrm)
    assert abs(np.linalg.norm(Vg[:, j]) - 1) < 1e-5
    assert np.abs(Vg[:, :j].T @ Vg[:, j]).max() < 1e-5


def test_dgks_second_pass_and_lindep(ks, ctx, oracle):
    rng = np.random.default_rng(5)
    n = 20000
    V = np.zeros((n, 4), order="F")
    V[:, :3] = np.linalg.qr(rng.standard_normal((n, 3)))[0]
    V[:, 3] = V[:, :3] @ np.array([1.0, -2.0, 0.5]) + 1e-6 * rng.standard_normal(n)
    r = _orth_case(ks, ctx, oracle, V, 3)
    assert r["passes"] == 2 and r["st"].total_passes == 2 and not r["st"].lindep
    assert np.allclose(r["h"][:4], r["H"], rtol=1e-8, atol=1e-15)
    assert np.abs(r["Vg"][:, :3].T @ r["Vg"][:, 3]).max() < 1e-14
    # the column the second pass stores, normalized in that pass's own sweep
    assert np.abs(r["Vg"][:, 3] - r["Vo"][:, 3]).max() < 1e-12
    assert abs(np.linalg.norm(r["Vg"][:, 3]) - 1) < 1e-13
    V[:, 3] = 0.0
    r = _orth_case(ks, ctx, oracle, V, 3)
    assert r["lindep"] and r["st"].lindep == 1 and r["st"].breakdown_col == 3 and r["st"].active == 0
    assert r["st"].nrm == 0.0 and r["h"][3] == 0.0


@pytest.mark.parametrize("refine", [1, 2])
def test_orthonormalize_refine_never_always(ks, ctx, oracle, refine):
    rng = np.random.default_rng(8)
    n, j = 4096, 6
    V = np.zeros((n, j + 1), order="F")
    V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
    V[:, j] = rng.standard_normal(n)
    r = _orth_case(ks, ctx, oracle, V, j, refine=refine)
    assert r["st"].total_passes == r["passes"]
    assert abs(r["st"].nrm - r["nrm"]) <= 1e-13 * r["nrm"]
    assert np.allclose(r["h"][:j + 1], r["H"], rtol=1e-10, atol=1e-13)
    assert np.abs(r["Vg"][:, j] - r["Vo"][:, j]).max() < 1e-12
    assert abs(np.linalg.norm(r["Vg"][:, j]) - 1) < 1e-13


def test_dgks_third_pass_vs_oracle(ks, ctx, oracle):
    # eta = 1 and a column 1e-12 off the span: the second pass's coefficients are
    # rounding-size but still shorten the vector, so DGKS takes its third pass;
    # the stored column, h and the norm against the oracle's.  The column is a
    # 1e-12 remainder of an O(1) vector, so two correct implementations that add
    # in different orders agree to about u |v| / |w''| (1e-5 relative) only
    rng = np.random.default_rng(9)
    n, j = 30011, 20
    V = np.zeros((n, j + 1), order="F")
    V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
    V[:, j] = V[:, :j] @ rng.standard_normal(j) + 1e-12 * rng.standard_normal(n)
    Vo = np.asfortranarray(V.copy())
    H, nrm, lindep, passes = oracle.orthonormalize_column(Vo, j, eta=1.0)
    dV = ks.DeviceArray.from_numpy(ctx, V)
    buf = ks.DeviceArray(ctx, (j + 2, j + 2))
    ctx.status_reset()
    ks.orthonormalize_column(ctx, ks.F64, n, dV, n, j, buf, j + 2, j, j, eta=1.0)
    st = ctx.status()
    Vg = dV.numpy()
    print(f"passes device={st.total_passes} oracle={passes} nrm={st.nrm:.3e}/{nrm:.3e}", flush=True)
    assert passes == 3 and st.total_passes == 3 and not st.lindep and not lindep
    assert abs(st.nrm - nrm) <= 1e-4 * nrm
    assert np.allclose(buf.numpy()[:j + 1, j], H, rtol=1e-8, atol=1e-15)
    assert np.abs(Vg[:, j] - Vo[:, j]).max() < 1e-4 * np.abs(Vo[:, j]).max()
    assert abs(np.linalg.norm(Vg[:, j]) - 1) < 1e-13
    assert np.abs(Vg[:, :j].T @ Vg[:, j]).max() < 1e-14


def test_dgks_second_pass_fp32(ks, ctx, oracle):
    # FP32 basis, a column 1e-3 off the span: two passes, each storing FP32; the
    # stored column against the FP64 oracle on the same FP32-representable data
    rng = np.random.default_rng(10)
    n, j = 20011, 12
    V = np.zeros((n, j + 1), order="F")
    V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
    V[:, j] = V[:, :j] @ rng.standard_normal(j) + 1e-3 * rng.standard_normal(n) / np.sqrt(n)
    V = np.asfortranarray(V.astype(np.float32).astype(np.float64))
    Vo = np.asfortranarray(V.copy())
    H, nrm, lindep, passes = oracle.orthonormalize_column(Vo, j)
    dV = ks.DeviceArray.from_numpy(ctx, np.asfortranarray(V.astype(np.float32)))
    buf = ks.DeviceArray(ctx, (j + 2, j + 2))
    ctx.status_reset()
    ks.orthonormalize_column(ctx, ks.F32, n, dV, n, j, buf, j + 2, j, j)
    st = ctx.status()
    Vg = dV.numpy().astype(np.float64)
    assert passes == 2 and st.total_passes == 2 and st.done == 1 and not st.lindep
    assert abs(st.nrm - nrm) <= 1e-3 * nrm
    assert abs(np.linalg.norm(Vg[:, j]) - 1) < 1e-6
    assert np.abs(Vg[:, :j].T @ Vg[:, j]).max() < 1e-6
    # the first pass's FP32 update cancels |v| / |w'| = 3.5e3 of the vector, so the
    # stored column carries u32 * 3.5e3 = 2e-4 relative error per operation
    assert np.abs(Vg[:, j] - Vo[:, j]).max() < 1e-2 * np.abs(Vo[:, j]).max()


@pytest.mark.parametrize("j,eta", [(12, 0.7071), (40, 0.7071), (12, 1.0), (40, 1.0), (60, 1.0)])
def test_lean_orthonormalize_halts_and_resumes(ks, ctx, oracle, j, eta):
    # the lean sequence enqueues two DGKS passes and no scaling.  A column with a
    # tiny component outside the span takes two passes at the default eta; at
    # eta = 1 the second pass's rounding-size coefficients also fall short of
    # the criterion, so the lean sequence stops after two with the status
    # naming the column and the resume runs the third.  Either way the result
    # is bit for bit that of the full sequence: V, the coefficient column, the
    # norm, lindep and the pass count (the oracle's)
    rng = np.random.default_rng(300 + j)
    n = 40961
    V = np.zeros((n, j + 1), order="F")
    V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
    V[:, j] = 1e-12 * rng.standard_normal(n) + V[:, :j] @ rng.standard_normal(j)
    ldbuf = j + 2

    def run(lean):
        dV = ks.DeviceArray.from_numpy(ctx, V)
        buf = ks.DeviceArray.from_numpy(ctx, np.full((ldbuf, j + 2), 3.0))
        ctx.status_reset()
        f = ks.lib().ks_bv_orthonormalize_column_lean if lean else ks.lib().ks_bv_orthonormalize_column
        ks.check(f(ctx.h, ks.F64, n, dV.ptr, n, j, buf.ptr, ldbuf, j, j, 0, eta, 1, 0), "orthonormalize")
        st = ctx.status()
        halted = (st.halt_col, st.halt_need)
        if lean and st.halt_col >= 0:
            ks.check(ks.lib().ks_bv_orthonormalize_resume(ctx.h, ks.F64, n, dV.ptr, n, j, buf.ptr, ldbuf, j, j, 0,
                                                          eta), "resume")
            st = ctx.status()
        return dV.numpy(), buf.numpy(), st, halted

    Va, ba, sa, _ = run(False)
    Vb, bb, sb, halted = run(True)
    H, nrm, lindep, passes = oracle.orthonormalize_column(np.asfortranarray(V.copy()), j, eta=eta)
    assert sa.total_passes == passes and sa.lindep == int(lindep)
    if eta == 1.0:
        assert passes == 3 and halted == (j, 1)
    else:
        assert passes == 2 and halted[0] == -1
    assert np.array_equal(Va, Vb) and np.array_equal(ba, bb)
    assert sb.total_passes == sa.total_passes and sb.nrm == sa.nrm and sb.lindep == sa.lindep
    assert sb.halt_col == -1 and sb.active == sa.active and sb.breakdown_col == sa.breakdown_col


def test_random_matches_oracle_bitwise(ks, ctx, oracle):
    n = 100003
    d = ks.DeviceArray(ctx, (n,))
    ks.random(ctx, ks.F64, n, d, 12345)
    assert np.array_equal(d.numpy(), oracle.random_uniform(12345, n))


def test_norm_scale_axpy(ks, ctx):
    rng = np.random.default_rng(3)
    for n in (1, 255, 257, 1000003, 16 * 1024 * 1024 + 5):
        x = rng.standard_normal(n)
        dx = ks.DeviceArray.from_numpy(ctx, x)
        assert abs(ks.norm(ctx, ks.F64, n, dx) - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
    x = rng.standard_normal(1000)
    y = rng.standard_normal(1000)
    dx = ks.DeviceArray.from_numpy(ctx, x)
    dy = ks.DeviceArray.from_numpy(ctx, y)
    ks.axpy(ctx, ks.F64, 1000, -3.0, dx, dy)
    ks.scale(ctx, ks.F64, 1000, dy, 0.5)
    assert np.allclose(dy.numpy(), 0.5 * (y - 3.0 * x), rtol=1e-15, atol=1e-15)
    ks.scale(ctx, ks.F64, 1000, dy, 0.0)
    assert not dy.numpy().any()


def test_grouped_step_second_pass_matches_alone(ks, oracle):
    # three solves in one grouped Arnoldi step, each column A v 1e-6 off the span of
    # its basis (A = I + 1e-6 L), so every member takes the second DGKS pass inside
    # the grouped launches; each member's stored column, h and pass count against the
    # same step done alone on its own context
    from slepc_amd import matgen as M
    specs = [(5003, 8), (20011, 20), (70001, 40)]
    ctxs, mats, Vs, bufs, alone = [], [], [], [], []
    grp = ks.Group()
    members = []
    for idx, (n, k) in enumerate(specs):
        rng = np.random.default_rng(n)
        csr = M.er_laplacian(n, 8.0, 7 + idx)
        rp, ci, va = csr
        val = 1e-6 * np.asarray(va, dtype=np.float64)
        for i in range(n):                      # + I on the diagonal (present in a Laplacian)
            for q in range(rp[i], rp[i + 1]):
                if ci[q] == i:
                    val[q] += 1.0
        ctx = ks.Context(0)
        A = ks.CSR(ctx, (rp, ci, val))
        V = np.zeros((n, k + 1), order="F")
        V[:, :k] = np.linalg.qr(rng.standard_normal((n, k)))[0]
        dV = ks.DeviceArray.from_numpy(ctx, V)
        buf = ks.DeviceArray(ctx, (k + 2, k + 2))
        # alone: SpMV of column k-1 into column k, then the orthonormalization
        dVa = ks.DeviceArray.from_numpy(ctx, V)
        bufa = ks.DeviceArray(ctx, (k + 2, k + 2))
        ctx.status_reset()
        check = ks.check
        check(ks.lib().ks_spmv(ctx.h, A.h, dVa.offset((k - 1) * n), dVa.offset(k * n), ks.F64, 0), "ks_spmv")
        ks.orthonormalize_column(ctx, ks.F64, n, dVa, n, k, bufa, k + 2, k, k)
        sta = ctx.status()
        alone.append((dVa.numpy()[:, k], bufa.numpy()[:k + 1, k], sta.total_passes, sta.nrm))
        ctx.status_reset()
        ctxs.append(ctx); mats.append(A); Vs.append(dV); bufs.append(buf)
        members.append(ks.GroupMember(ctx=ctx.h.value, A=A.h.value, V=dV.ptr.value, ld=n, n=n, k=k,
                                      buf=buf.ptr.value, ldbuf=k + 2, hcol=k, col=k))
        grp.wait_ctx(ctx)
    grp.step(members)
    for idx, (n, k) in enumerate(specs):
        grp.ctx_wait(ctxs[idx])
        st = ctxs[idx].status()
        col = Vs[idx].numpy()[:, k]
        h = bufs[idx].numpy()[:k + 1, k]
        col_a, h_a, passes_a, nrm_a = alone[idx]
        print(f"member n={n} k={k}: passes {st.total_passes}/{passes_a} nrm {st.nrm:.6e}/{nrm_a:.6e}", flush=True)
        assert st.total_passes == passes_a == 2 and st.done == 1 and not st.lindep
        assert abs(st.nrm - nrm_a) <= 1e-9 * nrm_a
        assert np.allclose(h, h_a, rtol=1e-10, atol=1e-18)
        assert np.abs(col - col_a).max() < 1e-9 * np.abs(col_a).max()
        assert abs(np.linalg.norm(col) - 1) < 1e-13
    grp.close()
    for c in ctxs:
        c.close()
