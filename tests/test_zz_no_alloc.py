"""The allocation-free column orthonormalization (ks_bv_orthonormalize_column_noalloc)
against the context-based entry point it shares its kernels with, through the C ABI:
same orthonormalized column, coefficient buffer and DGKS status bit for bit, on
consecutive calls that reuse the library's one scratch array (the second call
smaller than the first, then a linearly dependent column), and the size limit."""
import ctypes as C

import numpy as np
import pytest

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


def _basis(n, j, seed, dependent=False):
    rng = np.random.default_rng(seed)
    V = np.zeros((n, j + 1), order="F")
    V[:, :j] = np.linalg.qr(rng.standard_normal((n, j)))[0]
    if dependent:
        V[:, j] = 0.0
    else:
        # a large component in the span: DGKS takes a second pass
        V[:, j] = 1e-3 * rng.standard_normal(n) + V[:, :j] @ rng.standard_normal(j)
    return V


def _run(ks, ctx, V, j, prec, noalloc):
    n = V.shape[0]
    dt = np.float64 if prec == ks.F64 else np.float32
    dV = ks.DeviceArray.from_numpy(ctx, np.asfortranarray(V.astype(dt)))
    ldbuf = j + 2
    buf = ks.DeviceArray.from_numpy(ctx, np.full((ldbuf, j + 2), 7.0))
    if noalloc:
        dst = ks.DeviceArray(ctx, (C.sizeof(ks.Status),), np.uint8)
        print(f"noalloc: n={n} k={j} prec={prec}", flush=True)
        ks.orthonormalize_column_noalloc(ks.lib().ks_ctx_stream(ctx.h), prec, n, dV, n, j, buf, ldbuf, j, j,
                                         status_out=dst)
        ctx.sync()
        st = ks.Status.from_buffer_copy(dst.numpy().tobytes())
    else:
        ctx.status_reset()
        ks.orthonormalize_column(ctx, prec, n, dV, n, j, buf, ldbuf, j, j)
        st = ctx.status()
    return dV.numpy(), buf.numpy(), st


def _same(ks, ctx, V, j, prec):
    Va, ba, sa = _run(ks, ctx, V, j, prec, noalloc=False)
    Vb, bb, sb = _run(ks, ctx, V, j, prec, noalloc=True)
    assert np.array_equal(Va, Vb)
    assert np.array_equal(ba, bb)
    for f in ("done", "lindep", "total_passes", "breakdown_col", "active", "error"):
        assert getattr(sa, f) == getattr(sb, f), f
    assert sa.nrm == sb.nrm and sa.beta_last == sb.beta_last
    return sa


def test_noalloc_matches_context_path(ks, ctx):
    # call 1: the column-split sweep (48 columns, 6144 workgroups of partial sums)
    st = _same(ks, ctx, _basis(300007, 48, 1), 48, ks.F64)
    assert st.total_passes == 2 and st.done == 1 and not st.lindep
    # call 2, smaller: the register sweep (12 columns), FP64 then an FP32 basis
    st = _same(ks, ctx, _basis(10007, 12, 2), 12, ks.F64)
    assert st.total_passes == 2 and st.done == 1
    _same(ks, ctx, _basis(5003, 30, 3), 30, ks.F32)
    # a zero column after them: lindep, norm 0, nothing left over from the calls before
    st = _same(ks, ctx, _basis(4001, 6, 4, dependent=True), 6, ks.F64)
    assert st.lindep == 1 and st.breakdown_col == 6


def test_noalloc_size_limit(ks, ctx):
    n, j = 1000, 129
    dV = ks.DeviceArray(ctx, (n, j + 1))
    buf = ks.DeviceArray(ctx, (j + 2, j + 2))
    code = ks.lib().ks_bv_orthonormalize_column_noalloc(ks.lib().ks_ctx_stream(ctx.h), ks.F64, n, dV.ptr, n, j,
                                                        buf.ptr, j + 2, j, j, 0, 0.7071, 1, 0, None)
    assert code == -6   # KS_ERR_RANGE, refused on the host before any launch
