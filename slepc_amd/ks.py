"""ctypes binding of libks_hip.so (include/ks_hip.h): the C ABI of the
MI355X kernels.  Used by the kernel-level parity tests and the benchmark;
solvers go through the SLEPc-shaped layer in ``slepc_amd.eps``.

There is no fallback: loading fails loudly when the library is not built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libks_hip.so")

F64, F32 = 0, 1
REFINE_IFNEEDED, REFINE_NEVER, REFINE_ALWAYS = 0, 1, 2
KS_MAX_COLS = 320
PROF_CLASSES = ["spmv", "gs_dot_first", "gs_update_first", "gs_finish", "gs_update_refine", "scale",
                "norm", "axpy", "copy", "multinplace", "dotvec", "multvec", "other"]

_lib = None


class KSError(RuntimeError):
    pass


class Status(C.Structure):
    _fields_ = [("active", C.c_int32), ("breakdown_col", C.c_int32), ("col", C.c_int32),
                ("pass_done", C.c_int32), ("done", C.c_int32), ("lindep", C.c_int32),
                ("error", C.c_int32), ("need_norm", C.c_int32), ("total_passes", C.c_int64),
                ("onrm", C.c_double), ("nrm", C.c_double), ("beta_last", C.c_double),
                ("scale", C.c_double), ("sumsq", C.c_double), ("exact_sq", C.c_double),
                ("prescale", C.c_double), ("prescale_col", C.c_int32), ("prescale_pass", C.c_int32),
                ("halt_col", C.c_int32), ("halt_need", C.c_int32), ("defer_col", C.c_int32),
                ("defer_pass", C.c_int32)]


class GroupMember(C.Structure):
    """ks_group_member: one solve's column step inside a grouped launch."""
    _fields_ = [("ctx", C.c_void_p), ("A", C.c_void_p), ("V", C.c_void_p), ("ld", C.c_int), ("n", C.c_int),
                ("k", C.c_int), ("buf", C.c_void_p), ("ldbuf", C.c_int), ("hcol", C.c_int), ("col", C.c_int)]


class ProfEntry(C.Structure):
    _fields_ = [("launches", C.c_int64), ("skipped", C.c_int64), ("ms", C.c_double),
                ("bytes", C.c_double), ("flops", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KSError(f"{LIB_PATH} is not built; run `python -m slepc_amd.build`")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        vp, i32, i64, dbl, sz = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
        sigs = {
            "ks_init": [i32], "ks_ctx_create": [C.POINTER(vp)], "ks_ctx_destroy": [vp],
            "ks_ctx_sync": [vp], "ks_malloc": [C.POINTER(vp), sz], "ks_free": [vp],
            "ks_memcpy_htod": [vp, vp, vp, sz], "ks_memcpy_dtoh": [vp, vp, vp, sz],
            "ks_memcpy_dtod": [vp, vp, vp, sz], "ks_memset_zero": [vp, vp, sz],
            "ks_csr_create": [C.POINTER(vp), vp, i32, i32, i64, vp, vp, vp, i32],
            "ks_csr_destroy": [vp],
            "ks_csr_info": [vp, C.POINTER(i32), C.POINTER(i64), C.POINTER(i32), C.POINTER(i32),
                            C.POINTER(i32)],
            "ks_spmv": [vp, vp, vp, vp, i32, i32],
            "ks_bv_dotvec": [vp, i32, i32, vp, i32, i32, vp, vp],
            "ks_bv_multvec": [vp, i32, i32, dbl, vp, i32, i32, vp, dbl, vp],
            "ks_bv_orthonormalize_column": [vp, i32, i32, vp, i32, i32, vp, i32, i32, i32, i32, dbl,
                                            i32, i32],
            "ks_bv_orthonormalize_column_lean": [vp, i32, i32, vp, i32, i32, vp, i32, i32, i32, i32, dbl, i32,
                                                 i32],
            "ks_bv_orthonormalize_resume": [vp, i32, i32, vp, i32, i32, vp, i32, i32, i32, i32, dbl],
            "ks_bv_orthonormalize_column_noalloc": [vp, i32, i32, vp, i32, i32, vp, i32, i32, i32, i32, dbl,
                                                    i32, i32, vp],
            "ks_status_reset": [vp], "ks_status_get": [vp, C.POINTER(Status)],
            "ks_bv_scale": [vp, i32, i32, vp, dbl], "ks_bv_norm": [vp, i32, i32, vp, C.POINTER(dbl)],
            "ks_bv_axpy": [vp, i32, i32, dbl, vp, vp], "ks_bv_copy": [vp, i32, i32, vp, vp],
            "ks_bv_random": [vp, i32, i32, vp, C.c_uint64],
            "ks_bv_convert": [vp, i32, i32, vp, i32, vp],
            "ks_bv_multinplace": [vp, i32, i32, vp, i32, i32, vp, i32, i32, i32],
            "ks_prof_enable": [vp, i32], "ks_prof_enable_classes": [vp, C.c_uint], "ks_prof_collect": [vp],
            "ks_prof_get": [vp, i32, C.POINTER(ProfEntry)], "ks_prof_reset": [vp],
            "ks_mem_info": [C.POINTER(sz), C.POINTER(sz)],
            "ks_group_create": [C.POINTER(vp)], "ks_group_destroy": [vp],
            "ks_group_step": [vp, i32, vp, i32, dbl], "ks_group_wait_ctx": [vp, vp],
            "ks_ctx_wait_group": [vp, vp],
            "ks_lu_create": [vp, vp, dbl, C.POINTER(vp)], "ks_lu_solve": [vp, vp, vp, vp], "ks_lu_destroy": [vp],
            "ks_lu_check": [vp, vp],
            "ks_csr_transpose": [vp, vp, C.POINTER(vp)], "ks_csr_download": [vp, vp, vp, vp, vp],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.ks_error_string.argtypes = [C.c_int]
        L.ks_error_string.restype = C.c_char_p
        L.ks_ctx_stream.argtypes = [vp]
        L.ks_ctx_stream.restype = vp
        _lib = L
    return _lib


def check(code: int, what: str):
    if code != 0:
        msg = lib().ks_error_string(code).decode()
        raise KSError(f"{what} failed: {msg} ({code})")


def symbols():
    """Names of every entry point declared in include/ks_hip.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "ks_hip.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:int|void \*|const char \*)\s*(ks_\w+)\s*\(", text, re.M)))


class Context:
    def __init__(self, device: int = 0):
        L = lib()
        check(L.ks_init(device), "ks_init")
        h = C.c_void_p()
        check(L.ks_ctx_create(C.byref(h)), "ks_ctx_create")
        self.h = h

    def sync(self):
        check(lib().ks_ctx_sync(self.h), "ks_ctx_sync")

    def close(self):
        if self.h:
            lib().ks_ctx_destroy(self.h)
            self.h = None

    def status(self) -> Status:
        st = Status()
        check(lib().ks_status_get(self.h, C.byref(st)), "ks_status_get")
        return st

    def status_reset(self):
        check(lib().ks_status_reset(self.h), "ks_status_reset")

    # profiling
    def prof(self, on: bool):
        check(lib().ks_prof_enable(self.h, int(on)), "ks_prof_enable")

    def prof_reset(self):
        check(lib().ks_prof_reset(self.h), "ks_prof_reset")

    def prof_get(self) -> dict:
        check(lib().ks_prof_collect(self.h), "ks_prof_collect")
        out = {}
        for i, name in enumerate(PROF_CLASSES):
            e = ProfEntry()
            check(lib().ks_prof_get(self.h, i, C.byref(e)), "ks_prof_get")
            out[name] = dict(launches=e.launches, skipped=e.skipped, ms=e.ms, bytes=e.bytes, flops=e.flops)
        return out


class DeviceArray:
    """A flat device allocation holding a numpy-shaped array (C or F order)."""

    def __init__(self, ctx: Context, shape, dtype=np.float64, order="F"):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        self.dtype = np.dtype(dtype)
        self.order = order
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().ks_malloc(C.byref(p), max(self.nbytes, 8)), "ks_malloc")
        self.ptr = p

    @classmethod
    def from_numpy(cls, ctx, a, order="F"):
        a = np.asarray(a)
        d = cls(ctx, a.shape, a.dtype, order)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.array(a, dtype=self.dtype, order=self.order, copy=True)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            check(lib().ks_memcpy_htod(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes),
                  "ks_memcpy_htod")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype, order=self.order)
        if self.nbytes:
            check(lib().ks_memcpy_dtoh(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes),
                  "ks_memcpy_dtoh")
        return out

    def offset(self, elems: int) -> C.c_void_p:
        return C.c_void_p(self.ptr.value + int(elems) * self.dtype.itemsize)

    def free(self):
        if self.ptr:
            lib().ks_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CSR:
    def __init__(self, ctx: Context, csr, ncols: int | None = None, prec: int = F64):
        rowptr, colidx, val = csr
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colidx = np.ascontiguousarray(colidx, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64 if prec == F64 else np.float32)
        self.n = len(self.rowptr) - 1
        self.ncols = self.n if ncols is None else ncols
        self.nnz = int(self.rowptr[-1])
        self.prec = prec
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().ks_csr_create(C.byref(h), ctx.h, self.n, self.ncols, self.nnz,
                                  self.rowptr.ctypes.data_as(C.c_void_p),
                                  self.colidx.ctypes.data_as(C.c_void_p),
                                  self.val.ctypes.data_as(C.c_void_p), prec), "ks_csr_create")
        self.h = h

    def info(self):
        n, nnz, mr, var, nb = C.c_int(), C.c_int64(), C.c_int(), C.c_int(), C.c_int()
        check(lib().ks_csr_info(self.h, C.byref(n), C.byref(nnz), C.byref(mr), C.byref(var), C.byref(nb)),
              "ks_csr_info")
        return dict(n=n.value, nnz=nnz.value, max_row=mr.value, variant=var.value, nblocks=nb.value)

    def transpose(self) -> "CSR":
        """A^T built on the device (ks_csr_transpose)."""
        h = C.c_void_p()
        check(lib().ks_csr_transpose(self.h, self.ctx.h, C.byref(h)), "ks_csr_transpose")
        T = CSR.__new__(CSR)
        T.ctx, T.h, T.prec = self.ctx, h, self.prec
        info = T.info()
        T.n, T.nnz, T.ncols = info["n"], info["nnz"], self.n
        T.rowptr = T.colidx = T.val = None
        return T

    def download(self):
        """(rowptr, colidx, val) read back from the device (ks_csr_download)."""
        rp = np.zeros(self.n + 1, dtype=np.int32)
        ci = np.zeros(max(self.nnz, 1), dtype=np.int32)
        va = np.zeros(max(self.nnz, 1), dtype=np.float64 if self.prec == F64 else np.float32)
        check(lib().ks_csr_download(self.h, self.ctx.h, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                    va.ctypes.data_as(C.c_void_p)), "ks_csr_download")
        return rp, ci[:self.nnz], va[:self.nnz]

    def free(self):
        if self.h:
            lib().ks_csr_destroy(self.h)
            self.h = None


class LU:
    """ks_lu: the dense LU factors (partial pivoting) of A - sigma I on the device."""

    def __init__(self, ctx: Context, A: CSR, sigma: float = 0.0):
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().ks_lu_create(ctx.h, A.h, C.c_double(sigma), C.byref(self.h)), "ks_lu_create")

    def solve(self, b: DeviceArray, x: DeviceArray):
        check(lib().ks_lu_solve(self.ctx.h, self.h, b.ptr, x.ptr), "ks_lu_solve")

    def check(self):
        """Waits and raises if a solve's bounded wait ran out (ks_lu_check)."""
        check(lib().ks_lu_check(self.ctx.h, self.h), "ks_lu_check")

    def close(self):
        if self.h:
            lib().ks_lu_destroy(self.h)
            self.h = None


class Group:
    """ks_group: one Arnoldi step of several solves per grouped launch."""

    def __init__(self):
        self.h = C.c_void_p()
        check(lib().ks_group_create(C.byref(self.h)), "ks_group_create")

    def step(self, members, refine=REFINE_IFNEEDED, eta=0.7071):
        arr = (GroupMember * len(members))(*members)
        check(lib().ks_group_step(self.h, len(members), C.cast(arr, C.c_void_p), refine, eta), "ks_group_step")

    def wait_ctx(self, ctx):
        check(lib().ks_group_wait_ctx(self.h, ctx.h), "ks_group_wait_ctx")

    def ctx_wait(self, ctx):
        check(lib().ks_ctx_wait_group(ctx.h, self.h), "ks_ctx_wait_group")

    def close(self):
        if self.h:
            lib().ks_group_destroy(self.h)
            self.h = None


def spmv(ctx, A: CSR, x: DeviceArray, y: DeviceArray, prec=F64, gate=0):
    check(lib().ks_spmv(ctx.h, A.h, x.ptr, y.ptr, prec, gate), "ks_spmv")


def dotvec(ctx, prec, n, V, ld, k, y, q, voff=0, yoff=0):
    check(lib().ks_bv_dotvec(ctx.h, prec, n, V.offset(voff), ld, k, y.offset(yoff), q.ptr), "ks_bv_dotvec")


def multvec(ctx, prec, n, alpha, V, ld, k, q, beta, y, voff=0, yoff=0):
    check(lib().ks_bv_multvec(ctx.h, prec, n, alpha, V.offset(voff), ld, k, q.ptr, beta, y.offset(yoff)),
          "ks_bv_multvec")


def orthonormalize_column(ctx, prec, n, V, ld, k, buf, ldbuf, hcol, col, refine=REFINE_IFNEEDED,
                          eta=0.7071, scale=1, safe=0):
    check(lib().ks_bv_orthonormalize_column(ctx.h, prec, n, V.ptr, ld, k, buf.ptr, ldbuf, hcol, col, refine,
                                            eta, scale, safe), "ks_bv_orthonormalize_column")


def orthonormalize_column_noalloc(stream, prec, n, V, ld, k, buf, ldbuf, hcol, col, refine=REFINE_IFNEEDED,
                                  eta=0.7071, scale=1, safe=0, status_out=None):
    """The allocation-free variant: `stream` is a raw hipStream_t (ks_ctx_stream of a
    context), the workspace is the library's own device array."""
    check(lib().ks_bv_orthonormalize_column_noalloc(stream, prec, n, V.ptr, ld, k, buf.ptr, ldbuf, hcol, col,
                                                    refine, eta, scale, safe,
                                                    status_out.ptr if status_out is not None else None),
          "ks_bv_orthonormalize_column_noalloc")


def multinplace(ctx, prec, n, V, ld, k, Q, ldq, s, e, voff=0, qoff=0):
    check(lib().ks_bv_multinplace(ctx.h, prec, n, V.offset(voff), ld, k, Q.offset(qoff), ldq, s, e),
          "ks_bv_multinplace")


def norm(ctx, prec, n, x, off=0) -> float:
    r = C.c_double()
    check(lib().ks_bv_norm(ctx.h, prec, n, x.offset(off), C.byref(r)), "ks_bv_norm")
    return r.value


def random(ctx, prec, n, x, seed, off=0):
    check(lib().ks_bv_random(ctx.h, prec, n, x.offset(off), seed), "ks_bv_random")


def scale(ctx, prec, n, x, alpha, off=0):
    check(lib().ks_bv_scale(ctx.h, prec, n, x.offset(off), alpha), "ks_bv_scale")


def axpy(ctx, prec, n, alpha, x, y):
    check(lib().ks_bv_axpy(ctx.h, prec, n, alpha, x.ptr, y.ptr), "ks_bv_axpy")
