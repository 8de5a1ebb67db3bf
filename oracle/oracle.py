"""ctypes wrapper of libks_oracle.so -- the CPU restatement of SLEPc's
Krylov-Schur path.

TEST INFRASTRUCTURE ONLY (the checker): imported by tests/, by
``__graft_entry__.smoke()`` and by bench.py's cpu_baseline leg, never by the
product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libks_oracle.so")

LARGEST_MAGNITUDE, SMALLEST_MAGNITUDE, LARGEST_REAL, SMALLEST_REAL = 1, 2, 3, 4
LARGEST_IMAGINARY, SMALLEST_IMAGINARY, TARGET_MAGNITUDE, TARGET_REAL = 5, 6, 7, 8
REFINE_IFNEEDED, REFINE_NEVER, REFINE_ALWAYS = 0, 1, 2
RITZ, HARMONIC = 0, 1


class Opts(C.Structure):
    _fields_ = [("nev", C.c_int), ("ncv", C.c_int), ("mpd", C.c_int), ("max_it", C.c_int),
                ("tol", C.c_double), ("which", C.c_int), ("target", C.c_double),
                ("hermitian", C.c_int), ("keep", C.c_double), ("lock", C.c_int),
                ("eta", C.c_double), ("refine", C.c_int), ("nthreads", C.c_int),
                ("restart_cap", C.c_int), ("extraction", C.c_int), ("twosided", C.c_int),
                ("st_sinvert", C.c_int), ("pc_none", C.c_int), ("ksp_rtol", C.c_double), ("ksp_bcgs", C.c_int),
                ("pc_ilu", C.c_int), ("ksp_lu", C.c_int)]


class Info(C.Structure):
    _fields_ = [("nconv", C.c_int), ("its", C.c_int), ("reason", C.c_int),
                ("nmatmult", C.c_long), ("gs_passes", C.c_long),
                ("t_solve", C.c_double), ("t_ds", C.c_double), ("ksp_its", C.c_long)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run `make -C oracle`)")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a, ct=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def _csr_args(csr):
    rowptr, colidx, val = csr
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colidx = np.ascontiguousarray(colidx, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return rowptr, colidx, val


def random_uniform(seed: int, n: int) -> np.ndarray:
    out = np.empty(n)
    lib().or_random_uniform(C.c_ulonglong(seed), C.c_int(n), _p(out))
    return out


def spmv(csr, x: np.ndarray) -> np.ndarray:
    rowptr, colidx, val = _csr_args(csr)
    n = len(rowptr) - 1
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(n)
    rc = lib().or_spmv(C.c_int(n), _p(rowptr, C.c_int), _p(colidx, C.c_int), _p(val), _p(x), _p(y))
    assert rc == 0
    return y


def orthonormalize_column(V: np.ndarray, j: int, nc: int = 0, refine: int = REFINE_IFNEEDED,
                          eta: float = 0.7071):
    """V is (n, ncols) float64 Fortran-ordered, modified in place."""
    assert V.flags.f_contiguous and V.dtype == np.float64
    n = V.shape[0]
    H = np.zeros(nc + j + 1)
    norm = C.c_double()
    lindep = C.c_int()
    npasses = C.c_int()
    rc = lib().or_orthonormalize_column(C.c_int(n), C.c_int(n), _p(V), C.c_int(nc), C.c_int(j),
                                        C.c_int(refine), C.c_double(eta), _p(H), C.byref(norm),
                                        C.byref(lindep), C.byref(npasses))
    if rc:
        raise RuntimeError(f"or_orthonormalize_column failed ({rc})")
    return H, norm.value, bool(lindep.value), npasses.value


def multinplace(V: np.ndarray, Q: np.ndarray, l: int, k: int, s: int, e: int, nc: int = 0):
    assert V.flags.f_contiguous and Q.flags.f_contiguous
    n = V.shape[0]
    rc = lib().or_multinplace(C.c_int(n), C.c_int(n), _p(V), C.c_int(nc), C.c_int(l), C.c_int(k),
                              _p(Q), C.c_int(Q.shape[0]), C.c_int(s), C.c_int(e))
    assert rc == 0


def ds_translate_harmonic(H: np.ndarray, tau: float, beta: float):
    """DSTranslateHarmonic of the leading n x n block of H ((n+1) x n, Hessenberg):
    returns (H + g b^T, g, gamma)."""
    n = H.shape[1]
    ld = n + 1
    A = np.zeros((ld, ld), order="F")
    A[:n + 1, :n] = H
    g = np.zeros(ld)
    gamma = C.c_double()
    rc = lib().or_ds_translate_harmonic(C.c_int(ld), C.c_int(n), _p(A), C.c_double(tau), C.c_double(beta), _p(g),
                                        C.byref(gamma))
    if rc:
        raise RuntimeError(f"or_ds_translate_harmonic failed ({rc})")
    return A[:n, :n].copy(), g[:n].copy(), gamma.value


BLOCK_GS, BLOCK_CHOL, BLOCK_TSQR, BLOCK_TSQRCHOL, BLOCK_SVQB = 0, 1, 2, 3, 4


def bv_orthogonalize(V: np.ndarray, l: int = 0, block: int = BLOCK_GS):
    """BVOrthogonalize of columns l.. of V (n, k), in place; returns R (k, k)
    with columns l.. filled."""
    assert V.flags.f_contiguous and V.dtype == np.float64
    n, k = V.shape
    R = np.zeros((k, k), order="F")
    rc = lib().or_bv_orthogonalize(C.c_int(n), C.c_int(n), _p(V), C.c_int(l), C.c_int(k), C.c_int(block),
                                   _p(R), C.c_int(k))
    if rc:
        raise RuntimeError(f"or_bv_orthogonalize failed ({rc})")
    return R


def bv_tsqr_split(V: np.ndarray, nblocks: int):
    """TSQR of V (n, k) over ``nblocks`` row blocks split as PETSc splits rows over
    processes, in place; returns R (k, k)."""
    assert V.flags.f_contiguous and V.dtype == np.float64
    n, k = V.shape
    R = np.zeros((k, k), order="F")
    rc = lib().or_bv_tsqr_split(C.c_int(n), C.c_int(n), _p(V), C.c_int(k), C.c_int(nblocks), _p(R), C.c_int(k))
    if rc:
        raise RuntimeError(f"or_bv_tsqr_split failed ({rc})")
    return R


def ilu0(csr, sigma: float = 0.0) -> np.ndarray:
    """ILU(0) factor values of A - sigma I on A's pattern (L strictly below the
    diagonal, U on and above)."""
    rowptr = np.ascontiguousarray(csr[0], dtype=np.int32)
    colidx = np.ascontiguousarray(csr[1], dtype=np.int32)
    val = np.ascontiguousarray(csr[2], dtype=np.float64)
    lu = np.zeros(len(val))
    ip = C.POINTER(C.c_int)
    rc = lib().or_ilu0(C.c_int(len(rowptr) - 1), rowptr.ctypes.data_as(ip), colidx.ctypes.data_as(ip), _p(val),
                       C.c_double(sigma), _p(lu))
    if rc:
        raise RuntimeError(f"or_ilu0 failed ({rc})")
    return lu


def lu(csr, sigma: float = 0.0):
    """Dense LU with partial pivoting of A - sigma I: (LU by columns, L strictly
    below the diagonal and U on and above; ipiv, 0-based interchanges)."""
    rowptr = np.ascontiguousarray(csr[0], dtype=np.int32)
    colidx = np.ascontiguousarray(csr[1], dtype=np.int32)
    val = np.ascontiguousarray(csr[2], dtype=np.float64)
    n = len(rowptr) - 1
    a = np.zeros((n, n), order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    rc = lib().or_lu(C.c_int(n), rowptr.ctypes.data_as(ip), colidx.ctypes.data_as(ip), _p(val), C.c_double(sigma),
                     _p(a), ipiv.ctypes.data_as(ip))
    if rc:
        raise RuntimeError(f"or_lu failed ({rc})")
    return a, ipiv


def bv_dot(X: np.ndarray, Y: np.ndarray, M: np.ndarray, lx: int = 0, ly: int = 0):
    """M(ly:, lx:) = Y(:, ly:)^T X(:, lx:) in place (BVDot with leading columns)."""
    assert X.flags.f_contiguous and Y.flags.f_contiguous and M.flags.f_contiguous
    n = X.shape[0]
    rc = lib().or_bv_dot(C.c_int(n), _p(X), C.c_int(n), C.c_int(lx), C.c_int(X.shape[1]), _p(Y), C.c_int(n),
                         C.c_int(ly), C.c_int(Y.shape[1]), _p(M), C.c_int(M.shape[0]))
    assert rc == 0


def ds_run(dstype: str, ld: int, n: int, l: int = 0, k: int = 0, compact: bool = False,
           extrarow: bool = False, state: int = 0, A=None, T=None, which: int = LARGEST_MAGNITUDE,
           jvec: int = -1):
    """DSSolve + DSSort + [DSUpdateExtraRow] + DSVectors; returns a dict."""
    t = 1 if dstype == "hep" else 2
    A = None if A is None else np.array(A, dtype=np.float64, order="F", copy=True)
    T = None if T is None else np.array(T, dtype=np.float64, order="F", copy=True)
    wr = np.zeros(ld)
    wi = np.zeros(ld)
    Q = np.zeros((ld, ld), order="F")
    X = np.zeros((ld, ld), order="F")
    rnorm = C.c_double(0.0)
    rc = lib().or_ds_run(C.c_int(t), C.c_int(ld), C.c_int(n), C.c_int(l), C.c_int(k),
                         C.c_int(int(compact)), C.c_int(int(extrarow)), C.c_int(state),
                         _p(A), _p(T), C.c_int(which), _p(wr), _p(wi), _p(Q), C.c_int(jvec),
                         C.byref(rnorm), _p(X))
    if rc:
        raise RuntimeError(f"or_ds_run failed ({rc})")
    return dict(wr=wr[:n], wi=wi[:n], A=A, T=T, Q=Q, X=X, rnorm=rnorm.value)


def eps_solve(csr, nev: int, ncv: int = 0, mpd: int = 0, tol: float = 1e-8,
              which: int = LARGEST_MAGNITUDE, hermitian: bool = True, target: float = 0.0,
              v0=None, deflation=None, max_it: int = 0, keep: float = 0.5, lock: bool = True,
              eta: float = 0.7071, refine: int = REFINE_IFNEEDED, nthreads: int = 0,
              restart_cap: int = 0, want_vectors: bool = True, extraction: int = 0, twosided: bool = False,
              sinvert: bool = False, pc_none: bool = False, ksp_rtol: float = 0.0, ksp: str = "cg",
              pc: str = "jacobi"):
    rowptr, colidx, val = _csr_args(csr)
    n = len(rowptr) - 1
    o = Opts(nev=nev, ncv=ncv, mpd=mpd, max_it=max_it, tol=tol, which=which, target=target,
             hermitian=int(hermitian), keep=keep, lock=int(lock), eta=eta, refine=refine,
             nthreads=nthreads, restart_cap=restart_cap, extraction=extraction, twosided=int(twosided),
             st_sinvert=int(sinvert), pc_none=int(pc_none), ksp_rtol=ksp_rtol, ksp_bcgs=int(ksp == "bcgs"),
             pc_ilu=int(pc == "ilu"), ksp_lu=int(ksp == "preonly"))
    ncv_eff = ncv if ncv > 0 else max(2 * nev, nev + 15)
    ncv_eff = min(ncv_eff, n) if ncv_eff > 0 else n
    size = max(ncv_eff, nev) + 2
    eigr = np.zeros(size)
    eigi = np.zeros(size)
    errest = np.zeros(size)
    err = np.zeros(size)
    X = np.zeros((n, size), order="F") if want_vectors else None
    v0a = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
    nds = 0
    defl = None
    if deflation is not None:
        defl = np.asfortranarray(deflation, dtype=np.float64)
        nds = defl.shape[1]
    info = Info()
    Y = np.zeros((n, size), order="F") if (want_vectors and twosided) else None
    rc = lib().or_eps_solve_twosided(C.c_int(n), _p(rowptr, C.c_int), _p(colidx, C.c_int), _p(val),
                                     C.byref(o), _p(v0a), C.c_int(nds), _p(defl), _p(eigr), _p(eigi),
                                     _p(errest), _p(err), _p(X), _p(Y), C.byref(info))
    if rc:
        raise RuntimeError(f"or_eps_solve failed ({rc})")
    k = info.nconv
    out = dict(nconv=k, its=info.its, reason=info.reason, eigr=eigr[:k].copy(),
               eigi=eigi[:k].copy(), errest=errest[:k].copy(), err=err[:k].copy(),
               X=None if X is None else X[:, :k].copy(), nmatmult=info.nmatmult,
               gs_passes=info.gs_passes, t_solve=info.t_solve, t_ds=info.t_ds, ksp_its=info.ksp_its)
    if Y is not None:
        out["Y"] = Y[:, :k].copy()
    return out
