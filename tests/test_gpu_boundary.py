"""The literal drop-in of slepc_solver(HX,EX,HE,NX,NE)
(module_slepc_solver.F90:12-97) through ctypes: a caller's full host matrix
in, the NE smallest-magnitude eigenpairs out (EPS_SMALLEST_MAGNITUDE, :65),
checked against LAPACK (numpy.linalg.eigh) on positive and indefinite
Hermitian matrices of both solver paths (dense up to N = 1024, the block
solver above)."""
import sys

import numpy as np
import pytest

from wtb_amd import hip

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bse_like(rng, n, lo, hi, coupling):
    """Hermitian, diagonally dominant like a BSE kernel: a spread diagonal
    plus a small dense coupling"""
    d = np.sort(rng.uniform(lo, hi, n))
    a = coupling * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    h = 0.5 * (a + a.conj().T)
    h[np.diag_indices(n)] = d
    return h


def smallest_magnitude(h, ne):
    w, v = np.linalg.eigh(h)
    order = np.argsort(np.abs(w), kind="stable")[:ne]
    return w[order], v[:, order]


@pytest.mark.parametrize("n,ne,lo,hi", [(300, 20, 1.0, 8.0), (300, 25, -4.0, 5.0), (2000, 16, 2.0, 9.0)])
def test_slepc_solver_compat_smallest_magnitude(n, ne, lo, hi):
    say("case slepc compat", n, ne, lo, hi)
    rng = np.random.default_rng(n + ne)
    h = bse_like(rng, n, lo, hi, 0.3)
    w, _ = smallest_magnitude(h, ne)
    EX, HE = hip.slepc_solver_compat(h, ne)
    say("max |EX - LAPACK| ", np.abs(EX - w).max(), "lowest", np.linalg.eigvalsh(h)[0])
    assert np.abs(EX - w).max() < 1e-9
    assert np.all(np.diff(np.abs(EX)) >= 0)
    r = np.linalg.norm(h @ HE - HE * EX, axis=0) / np.abs(EX)
    assert r.max() < 1e-8
    assert np.allclose(HE.conj().T @ HE, np.eye(ne), atol=1e-8)
    if lo < 0:
        # the set is interior: the algebraically lowest eigenvalues are not in it
        assert np.linalg.eigvalsh(h)[0] < EX.min() - 1.0


def test_slepc_solver_compat_refuses_large_indefinite():
    say("case slepc compat indefinite N = 2000")
    rng = np.random.default_rng(7)
    h = bse_like(rng, 2000, -3.0, 6.0, 0.3)
    with pytest.raises(hip.WTBError) as e:
        hip.slepc_solver_compat(h, 16)
    assert e.value.code == -7
