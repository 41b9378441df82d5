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


def _dense_cases():
    """(n, structure) pairs over every size class of the dense reduction:
    n = 1..224 is held in registers and LDS in blocks of 32 rows, 225 and up
    streams from global memory"""
    sizes = [1, 2, 3, 15, 16, 17, 31, 32, 33, 48, 63, 64, 65, 96, 97, 128, 144, 160, 161, 184, 192, 193, 208,
             216, 223, 224, 225, 260]
    return [(n, "dense") for n in sizes] + [(n, "blocks") for n in (40, 130, 216)] + [(n, "degenerate") for n in
                                                                                    (72, 200, 260)]


@pytest.mark.parametrize("n,kind", _dense_cases())
def test_dense_eigensolver_sizes(n, kind):
    """Every eigenpair of a host matrix through the dense solver (N <= 1024)
    against LAPACK: energies to 1e-11 relative to the spectral radius,
    residuals and orthonormality to 1e-11.  'blocks' is block diagonal (zero
    columns below the subdiagonal: identity reflectors); 'degenerate' has
    eigenvalues of multiplicity four, whose vectors come from inverse
    iteration followed by Gram-Schmidt inside each cluster (orthonormal to
    1e-10)."""
    say("case dense eigensolver", n, kind)
    rng = np.random.default_rng(1000 + n)
    if kind == "dense":
        h = bse_like(rng, n, 1.0, 6.0, 0.5)
    elif kind == "blocks":
        h = np.zeros((n, n), complex)
        cut = [0, n // 5, n // 2, n // 2 + 1, n]
        for a, b in zip(cut[:-1], cut[1:]):
            h[a:b, a:b] = bse_like(rng, b - a, 1.0, 6.0, 0.5)
    else:
        lev = np.repeat(rng.uniform(1.0, 6.0, n // 4), 4)
        q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        h = (q * lev[None, :]) @ q.conj().T
        h = 0.5 * (h + h.conj().T)
    w = np.linalg.eigvalsh(h)
    EX, HE = hip.slepc_solver_compat(h, n)
    order = np.argsort(EX)
    EX, HE = EX[order], HE[:, order]
    scale = np.abs(w).max()
    say("max |E - LAPACK| / |H|", np.abs(EX - w).max() / scale)
    assert np.abs(EX - w).max() < 1e-11 * scale
    r = np.linalg.norm(h @ HE - HE * EX, axis=0) / scale
    orth = np.abs(HE.conj().T @ HE - np.eye(n)).max()
    say("max residual / |H|", r.max(), "orthonormality", orth)
    assert r.max() < 1e-11
    assert orth < (1e-10 if kind == "degenerate" else 1e-11)
