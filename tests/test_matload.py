"""The PETSc binary matrix reader (MatLoad format) on the reference's own data
file and its golden output: src/eps/tutorials/ex4.c reads
share/slepc/datafiles/matrices/rdb200.petsc (a copy is the fixture
tests/golden/rdb200.petsc) and output/ex4_1.out lists the 4 largest-magnitude
eigenvalues to 5 decimals.  CPU only: the reader runs on the host and the solve
here is the oracle's."""
import os
import struct

import numpy as np
import pytest

from slepc_amd import matgen as M

HERE = os.path.dirname(os.path.abspath(__file__))
RDB200 = os.path.join(HERE, "golden", "rdb200.petsc")
EX4_GOLDEN = "-35.00752, -34.10419, -33.20131, -32.68111"


def _read(path):
    from slepc_amd import eps as E
    return E.read_petsc_binary(path)


def test_rdb200_structure():
    rowptr, colidx, val = _read(RDB200)
    assert len(rowptr) == 201 and rowptr[-1] == 1120          # README: 200 x 200, 1120 entries
    for i in range(200):
        c = colidx[rowptr[i]:rowptr[i + 1]]
        assert np.all(np.diff(c) > 0) and c.min() >= 0 and c.max() < 200
    A = M.csr_to_dense(rowptr, colidx, val)
    # the stored values are symmetric although ex4 solves it as a general (NHEP) problem;
    # its spectrum has double eigenvalues, of which a single-vector Krylov method finds one copy
    assert np.array_equal(A, A.T)
    assert A[0, 0] == -3.294 and A[0, 1] == 4.0 and A[1, 3] == 3.872


def test_rdb200_ex4_golden_with_oracle(oracle):
    csr = _read(RDB200)
    # ex4 -eps_nev 4 -terse: Krylov-Schur defaults (ncv = max(2 nev, nev + 15), tol 1e-8, LM)
    r = oracle.eps_solve(csr, nev=4, ncv=19, tol=1e-8, hermitian=False, v0=oracle.random_uniform(0, 200))
    assert r["reason"] == 1 and r["nconv"] >= 4
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == EX4_GOLDEN
    # every returned value is an eigenvalue of the dense matrix, and together they are the
    # four largest-magnitude distinct ones
    ev = np.linalg.eigvalsh(M.csr_to_dense(*csr))
    distinct = np.unique(np.round(ev, 8))
    top = distinct[np.argsort(-np.abs(distinct))][:4]
    assert np.allclose(np.sort(r["eigr"][:4]), np.sort(top), rtol=1e-9)


def _write(path, classid, M_, N_, rowlen, cols, vals):
    with open(path, "wb") as f:
        f.write(struct.pack(">4i", classid, M_, N_, len(cols)))
        f.write(struct.pack(f">{len(rowlen)}i", *rowlen))
        f.write(struct.pack(f">{len(cols)}i", *cols))
        f.write(struct.pack(f">{len(vals)}d", *vals))


def test_reader_sorts_rows_and_rejects_bad_files(tmp_path):
    from slepc_amd import eps as E
    p = str(tmp_path / "a.petsc")
    # unsorted columns inside a row are sorted together with their values
    _write(p, 1211216, 2, 2, [2, 1], [1, 0, 1], [5.0, 3.0, 7.0])
    rowptr, colidx, val = _read(p)
    assert list(rowptr) == [0, 2, 3] and list(colidx) == [0, 1, 1] and list(val) == [3.0, 5.0, 7.0]
    _write(p, 1211215, 2, 2, [2, 1], [1, 0, 1], [5.0, 3.0, 7.0])          # wrong class id
    with pytest.raises(E.SlepcError):
        _read(p)
    _write(p, 1211216, 2, 2, [2, 1], [1, 2, 1], [5.0, 3.0, 7.0])          # column out of range
    with pytest.raises(E.SlepcError):
        _read(p)
    data = open(RDB200, "rb").read()
    open(p, "wb").write(data[:-8])                                         # truncated
    with pytest.raises(E.SlepcError):
        _read(p)
    with pytest.raises(E.SlepcError):
        _read(str(tmp_path / "missing.petsc"))
