"""The reference's test10, test20 and test24 goldens through the device path
(src/eps/tests/test10.c, test20.c, test24.c; outputs test10_1.out,
test20_1.out, test24_1.out): deflation of a graph Laplacian's null vector,
re-solving one EPS object after enlarging ncv, and exact eigenvectors as
constraints."""
import ctypes as C

import numpy as np
import pytest

from slepc_amd import matgen as M

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


@pytest.fixture(scope="module")
def E():
    from slepc_amd import eps as E
    E.initialize(0)
    return E


def test_test10_fiedler(E):
    r = E.solve(M.grid_graph_laplacian(10, 11), nev=4, which=E.EPS.SMALLEST_REAL, max_it=500,
                deflation=np.ones((110, 1)))
    assert r["reason"] == 1
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == "0.08101, 0.09789, 0.17890, 0.31749"


def test_test20_resolve_with_larger_ncv(E):
    tol = max(1000 * EPS, 1e-9)
    A = E.Mat(M.laplacian_1d(18))
    eps = E.EPS()
    eps.setOperators(A)
    eps.setProblemType(E.EPS.HEP)
    eps.setTolerances(tol)
    eps.setWhichEigenpairs(E.EPS.SMALLEST_REAL)
    eps.solve()
    first = eps.getEigenvalue(0)
    nev, ncv = C.c_int(), C.c_int()
    E.lib().EPSGetDimensions(eps.h, C.byref(nev), C.byref(ncv), None)
    assert ncv.value == 16
    eps.setDimensions(nev.value, ncv.value + 2)
    eps.solve()
    second = eps.getEigenvalue(0)
    assert "%.5f" % first == "0.02728" and "%.5f" % second == "0.02728"
    assert eps.computeError(0) < 5 * tol
    eps.destroy()
    A.destroy()


def test_test24_constraints(E):
    n, ncon = 30, 2
    alpha, beta = np.pi / (n + 1), np.sqrt(2.0 / (n + 1))
    j = np.arange(1, n + 1)
    vc = np.column_stack([np.sin(alpha * j * (i + 1)) * beta for i in range(ncon)])
    r = E.solve(M.laplacian_1d(n), nev=4, tol=1e-8, max_it=1200, which=E.EPS.SMALLEST_REAL, deflation=vc)
    assert r["reason"] == 1
    assert ", ".join("%.5f" % v for v in r["eigr"][:4]) == "0.09172, 0.16208, 0.25131, 0.35847"
