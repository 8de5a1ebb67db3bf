"""Python front-end of libslepc_amd.so (include/slepc_amd.h), named after
slepc4py's EPS/BV/Mat/Vec so that tests read like the reference's.  Every call
goes through the C host framework, which drives the HIP kernels; there is no
CPU fallback and loading fails loudly when the libraries are not built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import ks as _ks

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libslepc_amd.so")

F64, F32 = 0, 1
_lib = None
_initialized = False


class SlepcError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        _ks.lib()   # load the kernel library first (RTLD_GLOBAL)
        if not os.path.exists(LIB_PATH):
            raise SlepcError(f"{LIB_PATH} is not built; run `python -m slepc_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.SlepcAMDGetLastError.restype = C.c_char_p
        L.SlepcAMDGetDefaultContext.restype = C.c_void_p
        _lib = L
    return _lib


def _chk(code, what):
    if code != 0:
        msg = lib().SlepcAMDGetLastError().decode(errors="replace")
        raise SlepcError(f"{what} failed with error {code}: {msg}")


def initialize(device: int = 0):
    global _initialized
    if not _initialized:
        _chk(lib().SlepcInitialize(C.c_int(device)), "SlepcInitialize")
        _initialized = True


def krylov_extension() -> bool:
    """Whether BVHIPKS offers its device Krylov expansion (the default; the
    options database's -bv_hipks_krylov 0, from PETSC_OPTIONS, turns it off)."""
    opts = os.environ.get("PETSC_OPTIONS", "").split()
    for i, o in enumerate(opts):
        if o == "-bv_hipks_krylov":
            v = opts[i + 1] if i + 1 < len(opts) and not opts[i + 1].startswith("-") else "1"
            return v.lower() not in ("0", "false", "no")
    return True


def context_handle() -> C.c_void_p:
    initialize()
    return C.c_void_p(lib().SlepcAMDGetDefaultContext())


class Vec:
    def __init__(self, n: int, prec: int = F64):
        initialize()
        self.h = C.c_void_p()
        self.n = n
        _chk(lib().VecCreateHIP(C.c_int(n), C.c_int(prec), C.byref(self.h)), "VecCreateHIP")

    @classmethod
    def from_array(cls, a, prec: int = F64):
        a = np.ascontiguousarray(a, dtype=np.float64)
        v = cls(len(a), prec)
        v.set(a)
        return v

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.n,)
        _chk(lib().VecSetValuesHost(self.h, a.ctypes.data_as(C.POINTER(C.c_double))), "VecSetValuesHost")

    @classmethod
    def random(cls, n: int, seed: int, prec: int = F64):
        v = cls(n, prec)
        _chk(lib().VecSetRandomSeeded(v.h, C.c_uint64(seed)), "VecSetRandomSeeded")
        return v

    def array(self) -> np.ndarray:
        out = np.empty(self.n)
        _chk(lib().VecGetValuesHost(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "VecGetValuesHost")
        return out

    def norm(self) -> float:
        r = C.c_double()
        _chk(lib().VecNorm(self.h, C.c_int(1), C.byref(r)), "VecNorm")
        return r.value

    def destroy(self):
        if self.h:
            lib().VecDestroy(C.byref(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BV:
    """A BVHIPKS basis of m columns of length n (BVCreate/BVSetSizes/BVSetType),
    with host-side column access for tests and small drivers."""
    GS, CHOL, TSQR, TSQRCHOL, SVQB = 0, 1, 2, 3, 4   # BVOrthogBlockType

    def __init__(self, n: int, m: int, prec: int = F64):
        initialize()
        self.n, self.m = n, m
        self.h = C.c_void_p()
        _chk(lib().BVCreate(C.byref(self.h)), "BVCreate")
        _chk(lib().BVSetSizes(self.h, C.c_int(n), C.c_int(n), C.c_int(m)), "BVSetSizes")
        if prec != F64:
            _chk(lib().BVHIPKSSetPrecision(self.h, C.c_int(prec)), "BVHIPKSSetPrecision")
        _chk(lib().BVSetType(self.h, b"hipks"), "BVSetType")

    def set_columns(self, X, start: int = 0):
        X = np.asarray(X, dtype=np.float64)
        for j in range(X.shape[1]):
            col = np.ascontiguousarray(X[:, j])
            _chk(lib().BVSetColumnValuesHost(self.h, C.c_int(start + j), col.ctypes.data_as(C.POINTER(C.c_double))),
                 "BVSetColumnValuesHost")

    def get_columns(self, start: int = 0, stop: int | None = None) -> np.ndarray:
        stop = self.m if stop is None else stop
        out = np.zeros((self.n, stop - start), order="F")
        for j in range(start, stop):
            col = np.zeros(self.n)
            _chk(lib().BVGetColumnValuesHost(self.h, C.c_int(j), col.ctypes.data_as(C.POINTER(C.c_double))),
                 "BVGetColumnValuesHost")
            out[:, j - start] = col
        return out

    def set_active(self, l: int, k: int):
        _chk(lib().BVSetActiveColumns(self.h, C.c_int(l), C.c_int(k)), "BVSetActiveColumns")

    def insert_constraints(self, Cm) -> int:
        """BVInsertConstraints with the columns of Cm (orthonormalized on insertion;
        the regular columns' contents are lost).  Returns how many were kept."""
        Cm = np.asarray(Cm, dtype=np.float64)
        vs = [Vec.from_array(Cm[:, i]) for i in range(Cm.shape[1])]
        nc = C.c_int(len(vs))
        arr = (C.c_void_p * len(vs))(*[v.h for v in vs])
        _chk(lib().BVInsertConstraints(self.h, C.byref(nc), arr), "BVInsertConstraints")
        for v in vs:
            v.destroy()
        return nc.value

    def get_constraints(self, nc: int) -> np.ndarray:
        """The constraint columns (BVGetColumn with negative indices)."""
        return self.get_columns(-nc, 0)

    def set_block(self, block: int):
        _chk(lib().BVSetOrthogonalization(self.h, C.c_int(0), C.c_int(0), C.c_double(0.7071), C.c_int(block)),
             "BVSetOrthogonalization")

    @staticmethod
    def _dense(m: int, n: int, M=None):
        arr = np.zeros((m, n), order="F") if M is None else np.asfortranarray(M, dtype=np.float64)
        h = C.c_void_p()
        _chk(lib().MatCreateSeqDense(C.c_int(m), C.c_int(n), arr.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h)),
             "MatCreateSeqDense")
        return arr, h

    def orthogonalize(self, want_r: bool = True) -> np.ndarray | None:
        """BVOrthogonalize on the active columns; returns R (m x m) or None."""
        if not want_r:
            _chk(lib().BVOrthogonalize(self.h, None), "BVOrthogonalize")
            return None
        R, h = self._dense(self.m, self.m)
        try:
            _chk(lib().BVOrthogonalize(self.h, h), "BVOrthogonalize")
        finally:
            lib().MatDestroy(C.byref(h))
        return R

    def dot(self, Y: "BV", M=None) -> np.ndarray:
        """BVDot(self, Y, M): M = Y^T self over the active columns (other entries kept)."""
        M, h = self._dense(Y.m, self.m, M)
        try:
            _chk(lib().BVDot(self.h, Y.h, h), "BVDot")
        finally:
            lib().MatDestroy(C.byref(h))
        return M

    def mat_project(self, A, Y: "BV", M=None) -> np.ndarray:
        """BVMatProject(self, A, Y, M): M = Y^T A self (A a Mat or None)."""
        M, h = self._dense(Y.m, self.m, M)
        try:
            _chk(lib().BVMatProject(self.h, A.h if A is not None else None, Y.h, h), "BVMatProject")
        finally:
            lib().MatDestroy(C.byref(h))
        return M

    def destroy(self):
        if self.h:
            lib().BVDestroy(C.byref(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def read_petsc_binary(path: str):
    """(rowptr, colidx, val) of a matrix stored with PETSc's binary viewer, read by
    MatReadPETScBinary (host only: no device needed)."""
    L = lib()
    m, n = C.c_int(), C.c_int()
    rp, ci, va = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    _chk(L.MatReadPETScBinary(os.fsencode(path), C.byref(m), C.byref(n), C.byref(rp), C.byref(ci), C.byref(va)),
         "MatReadPETScBinary")
    try:
        rowptr = np.ctypeslib.as_array(rp, shape=(m.value + 1,)).astype(np.int32)
        nnz = int(rowptr[-1])
        colidx = np.ctypeslib.as_array(ci, shape=(max(nnz, 1),))[:nnz].astype(np.int32)
        val = np.ctypeslib.as_array(va, shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        L.MatFreePETScBinary(rp, ci, va)
    if m.value != n.value:
        raise SlepcError(f"{path}: {m.value} x {n.value} matrix is not square")
    return rowptr, colidx, val


class Mat:
    """MATHIPCSR matrix from host CSR arrays (32-bit indices)."""

    def __init__(self, csr, prec: int = F64):
        initialize()
        rowptr, colidx, val = csr
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colidx = np.ascontiguousarray(colidx, dtype=np.int32)
        self.n = len(self.rowptr) - 1
        self.h = C.c_void_p()
        L = lib()
        _chk(L.MatCreate(C.byref(self.h)), "MatCreate")
        _chk(L.MatSetSizes(self.h, self.n, self.n, self.n, self.n), "MatSetSizes")
        _chk(L.MatSetType(self.h, b"hipcsr"), "MatSetType")
        ip = self.rowptr.ctypes.data_as(C.c_void_p)
        jp = self.colidx.ctypes.data_as(C.c_void_p)
        if prec == F64:
            v = np.ascontiguousarray(val, dtype=np.float64)
            _chk(L.MatSeqAIJSetPreallocationCSR(self.h, ip, jp, v.ctypes.data_as(C.c_void_p)),
                 "MatSeqAIJSetPreallocationCSR")
        else:
            v = np.ascontiguousarray(val, dtype=np.float32)
            _chk(L.MatHIPCSRSetArraysF32(self.h, ip, jp, v.ctypes.data_as(C.c_void_p)), "MatHIPCSRSetArraysF32")

    @classmethod
    def load(cls, path: str):
        """MatLoad of a PETSc binary matrix file into a device CSR matrix."""
        return cls(read_petsc_binary(path))

    def mult(self, x: Vec, y: Vec):
        _chk(lib().MatMult(self.h, x.h, y.h), "MatMult")

    def multTranspose(self, x: Vec, y: Vec):
        _chk(lib().MatMultTranspose(self.h, x.h, y.h), "MatMultTranspose")

    def info(self):
        nnz, mr, var, nb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _chk(lib().MatHIPCSRGetInfo(self.h, C.byref(nnz), C.byref(mr), C.byref(var), C.byref(nb)),
             "MatHIPCSRGetInfo")
        return dict(nnz=nnz.value, max_row=mr.value, variant=var.value, nblocks=nb.value)

    def destroy(self):
        if self.h:
            lib().MatDestroy(C.byref(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class EPS:
    # EPSProblemType
    HEP, NHEP = 1, 3
    # EPSWhich
    LARGEST_MAGNITUDE, SMALLEST_MAGNITUDE, LARGEST_REAL, SMALLEST_REAL = 1, 2, 3, 4
    LARGEST_IMAGINARY, SMALLEST_IMAGINARY, TARGET_MAGNITUDE, TARGET_REAL = 5, 6, 7, 8
    # EPSConvergedReason
    CONVERGED_TOL, DIVERGED_ITS, DIVERGED_BREAKDOWN = 1, -1, -2
    # EPSErrorType
    ERROR_ABSOLUTE, ERROR_RELATIVE = 0, 1
    REFINE_IFNEEDED, REFINE_NEVER, REFINE_ALWAYS = 0, 1, 2
    # EPSExtraction (Krylov-Schur: RITZ, HARMONIC)
    RITZ, HARMONIC = 0, 1

    def __init__(self):
        initialize()
        self.h = C.c_void_p()
        _chk(lib().EPSCreate(C.byref(self.h)), "EPSCreate")
        self._keep = []
        self.n = None

    # ---- settings
    def setOperators(self, A: Mat):
        self._A = A
        self.n = A.n
        _chk(lib().EPSSetOperators(self.h, A.h, None), "EPSSetOperators")

    def setProblemType(self, t):
        _chk(lib().EPSSetProblemType(self.h, C.c_int(t)), "EPSSetProblemType")

    def setWhichEigenpairs(self, w):
        _chk(lib().EPSSetWhichEigenpairs(self.h, C.c_int(w)), "EPSSetWhichEigenpairs")

    def setTarget(self, t: float):
        _chk(lib().EPSSetTarget(self.h, C.c_double(t)), "EPSSetTarget")

    def setExtraction(self, extr: int):
        _chk(lib().EPSSetExtraction(self.h, C.c_int(extr)), "EPSSetExtraction")

    def setST(self, st_type: str = "shift", shift=None, ksp_rtol=None, ksp_maxits=None, pc_type=None,
              ksp_type=None):
        """STSetType (+ STSetShift) and the inner solver of shift-and-invert
        (STGetKSP -> KSPSetType ("preonly", the default direct solve with
        "lu"; "cg" or "bcgs"), KSPSetTolerances, KSPGetPC -> PCSetType ("lu"
        with "preonly"; "jacobi", "none", or "ilu" with "bcgs"))."""
        st = C.c_void_p()
        _chk(lib().EPSGetST(self.h, C.byref(st)), "EPSGetST")
        _chk(lib().STSetType(st, st_type.encode()), "STSetType")
        if shift is not None:
            _chk(lib().STSetShift(st, C.c_double(shift)), "STSetShift")
        if st_type == "sinvert":
            ksp, pc = C.c_void_p(), C.c_void_p()
            _chk(lib().STGetKSP(st, C.byref(ksp)), "STGetKSP")
            if ksp_type is not None:
                _chk(lib().KSPSetType(ksp, ksp_type.encode()), "KSPSetType")
            _chk(lib().KSPSetTolerances(ksp, C.c_double(-1.0 if ksp_rtol is None else ksp_rtol), C.c_double(-1.0),
                                        C.c_double(-1.0), C.c_int(0 if ksp_maxits is None else ksp_maxits)),
                 "KSPSetTolerances")
            if pc_type is not None:
                _chk(lib().KSPGetPC(ksp, C.byref(pc)), "KSPGetPC")
                _chk(lib().PCSetType(pc, pc_type.encode()), "PCSetType")

    def kspIterations(self) -> int:
        st, ksp, its = C.c_void_p(), C.c_void_p(), C.c_int()
        _chk(lib().EPSGetST(self.h, C.byref(st)), "EPSGetST")
        _chk(lib().STGetKSP(st, C.byref(ksp)), "STGetKSP")
        _chk(lib().KSPGetTotalIterations(ksp, C.byref(its)), "KSPGetTotalIterations")
        return its.value

    def setTwoSided(self, flg: bool = True):
        _chk(lib().EPSSetTwoSided(self.h, C.c_int(int(flg))), "EPSSetTwoSided")

    def setDimensions(self, nev=-2, ncv=-1, mpd=-1):
        _chk(lib().EPSSetDimensions(self.h, C.c_int(nev), C.c_int(ncv), C.c_int(mpd)), "EPSSetDimensions")

    def setTolerances(self, tol=-1.0, max_it=-1):
        _chk(lib().EPSSetTolerances(self.h, C.c_double(tol), C.c_int(max_it)), "EPSSetTolerances")

    def setKrylovSchurRestart(self, keep: float):
        self._ensure_type()
        _chk(lib().EPSKrylovSchurSetRestart(self.h, C.c_double(keep)), "EPSKrylovSchurSetRestart")

    def setKrylovSchurLocking(self, lock: bool):
        self._ensure_type()
        _chk(lib().EPSKrylovSchurSetLocking(self.h, C.c_int(int(lock))), "EPSKrylovSchurSetLocking")

    def setMixedPrecision(self, tol32: float):
        """FP32-basis phase down to tol32, then the FP64 restarts (tol32 <= 0: off)."""
        self._ensure_type()
        _chk(lib().EPSKrylovSchurSetMixedPrecision(self.h, C.c_double(tol32)), "EPSKrylovSchurSetMixedPrecision")

    def mixedStats(self) -> dict:
        its, nc = C.c_int(), C.c_int()
        nm, gp = C.c_int64(), C.c_int64()
        t = C.c_double()
        _chk(lib().EPSKrylovSchurGetMixedStats(self.h, C.byref(its), C.byref(nc), C.byref(nm), C.byref(gp),
                                               C.byref(t)), "EPSKrylovSchurGetMixedStats")
        return dict(its32=its.value, nconv32=nc.value, nmatmult32=nm.value, gs_passes32=gp.value, t32=t.value)

    def _ensure_type(self):
        t = C.c_char_p()
        lib().EPSGetType.argtypes = [C.c_void_p, C.POINTER(C.c_char_p)]
        _chk(lib().EPSGetType(self.h, C.byref(t)), "EPSGetType")
        if not t.value:
            _chk(lib().EPSSetType(self.h, b"krylovschur"), "EPSSetType")

    def _bv(self):
        bv = C.c_void_p()
        _chk(lib().EPSGetBV(self.h, C.byref(bv)), "EPSGetBV")
        return bv

    def setBVPrecision(self, prec: int):
        _chk(lib().BVHIPKSSetPrecision(self._bv(), C.c_int(prec)), "BVHIPKSSetPrecision")

    def setStepGraphs(self, flg: bool = True):
        """Replay each Arnoldi step from a hipGraph captured once per column."""
        self._ensure_type()
        _chk(lib().EPSKrylovSchurSetStepGraphs(self.h, C.c_int(int(flg))), "EPSKrylovSchurSetStepGraphs")

    def setOrthogonalization(self, refine=0, eta=0.7071):
        _chk(lib().BVSetOrthogonalization(self._bv(), C.c_int(0), C.c_int(refine), C.c_double(eta), C.c_int(0)),
             "BVSetOrthogonalization")

    def setInitialSpace(self, vecs):
        vs = [v if isinstance(v, Vec) else Vec.from_array(v) for v in vecs]
        self._keep.append(vs)
        arr = (C.c_void_p * len(vs))(*[v.h for v in vs])
        _chk(lib().EPSSetInitialSpace(self.h, C.c_int(len(vs)), arr), "EPSSetInitialSpace")

    def setDeflationSpace(self, vecs):
        vs = [v if isinstance(v, Vec) else Vec.from_array(v) for v in vecs]
        self._keep.append(vs)
        arr = (C.c_void_p * len(vs))(*[v.h for v in vs])
        _chk(lib().EPSSetDeflationSpace(self.h, C.c_int(len(vs)), arr), "EPSSetDeflationSpace")

    # ---- solve and results
    def setUp(self):
        _chk(lib().EPSSetUp(self.h), "EPSSetUp")

    def solve(self):
        _chk(lib().EPSSolve(self.h), "EPSSolve")

    def getConverged(self) -> int:
        n = C.c_int()
        _chk(lib().EPSGetConverged(self.h, C.byref(n)), "EPSGetConverged")
        return n.value

    def getIterationNumber(self) -> int:
        n = C.c_int()
        _chk(lib().EPSGetIterationNumber(self.h, C.byref(n)), "EPSGetIterationNumber")
        return n.value

    def getConvergedReason(self) -> int:
        n = C.c_int()
        _chk(lib().EPSGetConvergedReason(self.h, C.byref(n)), "EPSGetConvergedReason")
        return n.value

    def getEigenvalue(self, i: int):
        re, im = C.c_double(), C.c_double()
        _chk(lib().EPSGetEigenvalue(self.h, C.c_int(i), C.byref(re), C.byref(im)), "EPSGetEigenvalue")
        return complex(re.value, im.value) if im.value != 0.0 else re.value

    def getEigenvector(self, i: int, imag: bool = False):
        vr = Vec(self.n)
        vi = Vec(self.n) if imag else None
        _chk(lib().EPSGetEigenvector(self.h, C.c_int(i), vr.h, vi.h if vi else None), "EPSGetEigenvector")
        return (vr.array(), vi.array()) if imag else vr.array()

    def getLeftEigenvector(self, i: int, imag: bool = False):
        wr = Vec(self.n)
        wi = Vec(self.n) if imag else None
        _chk(lib().EPSGetLeftEigenvector(self.h, C.c_int(i), wr.h, wi.h if wi else None), "EPSGetLeftEigenvector")
        return (wr.array(), wi.array()) if imag else wr.array()

    def getErrorEstimate(self, i: int) -> float:
        r = C.c_double()
        _chk(lib().EPSGetErrorEstimate(self.h, C.c_int(i), C.byref(r)), "EPSGetErrorEstimate")
        return r.value

    def computeError(self, i: int, etype: int = 1) -> float:
        r = C.c_double()
        _chk(lib().EPSComputeError(self.h, C.c_int(i), C.c_int(etype), C.byref(r)), "EPSComputeError")
        return r.value

    def stats(self) -> dict:
        nm, gp = C.c_int64(), C.c_int64()
        its = C.c_int()
        tds, ts = C.c_double(), C.c_double()
        _chk(lib().EPSGetSolveStats(self.h, C.byref(nm), C.byref(gp), C.byref(its), C.byref(tds), C.byref(ts)),
             "EPSGetSolveStats")
        return dict(nmatmult=nm.value, gs_passes=gp.value, its=its.value, t_ds=tds.value, t_solve=ts.value)

    def eigenvalues(self):
        k = self.getConverged()
        return np.array([self.getEigenvalue(i) for i in range(k)])

    def destroy(self):
        if self.h:
            lib().EPSDestroy(C.byref(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def solve(csr, nev, ncv=-1, tol=1e-8, which=EPS.LARGEST_MAGNITUDE, hermitian=True, v0=None, max_it=-1,
          prec=F64, deflation=None, lock=True, keep=0.5, want_vectors=True, mixed_tol=0.0, step_graphs=False,
          target=None, extraction=0, twosided=False, st=None):
    """One-call convenience wrapper: returns a dict like the oracle's eps_solve.
    FP32 values in csr[2] keep the operator in FP32; ``prec`` is the basis precision."""
    A = Mat(csr, prec=F32 if np.asarray(csr[2]).dtype == np.float32 else F64)
    eps = EPS()
    eps.setOperators(A)
    eps.setProblemType(EPS.HEP if hermitian else EPS.NHEP)
    eps.setDimensions(nev, ncv)
    eps.setTolerances(tol, max_it)
    eps.setWhichEigenpairs(which)
    if target is not None:
        eps.setTarget(target)
    if extraction:
        eps.setExtraction(extraction)
    if twosided:
        eps.setTwoSided(True)
    if st is not None:            # dict(st_type="sinvert", shift=..., ksp_type="cg", ksp_rtol=..., pc_type="jacobi")
        eps.setST(**st)
    eps.setKrylovSchurLocking(lock)
    eps.setKrylovSchurRestart(keep)
    if mixed_tol > 0:
        eps.setMixedPrecision(mixed_tol)
    if prec != F64:
        eps.setBVPrecision(prec)
    if step_graphs:
        eps.setStepGraphs(True)
    if v0 is not None:
        eps.setInitialSpace([v0])
    if deflation is not None:
        eps.setDeflationSpace([deflation[:, i] for i in range(deflation.shape[1])])
    eps.solve()
    k = eps.getConverged()
    vals = [eps.getEigenvalue(i) for i in range(k)]
    out = dict(nconv=k, its=eps.getIterationNumber(), reason=eps.getConvergedReason(),
               eigr=np.array([np.real(v) for v in vals]), eigi=np.array([np.imag(v) for v in vals]),
               errest=np.array([eps.getErrorEstimate(i) for i in range(k)]),
               err=np.array([eps.computeError(i) for i in range(k)]))
    if want_vectors:
        X = np.zeros((A.n, k), order="F")
        for i in range(k):
            X[:, i] = eps.getEigenvector(i)
        out["X"] = X
        if twosided:
            Y = np.zeros((A.n, k), order="F")
            for i in range(k):
                Y[:, i] = eps.getLeftEigenvector(i)
            out["Y"] = Y
    out.update(eps.stats())
    if st is not None and st.get("st_type") == "sinvert":
        out["ksp_its"] = eps.kspIterations()
    if mixed_tol > 0:
        out.update(eps.mixedStats())
    eps.destroy()
    A.destroy()
    return out
