"""Finite-Q exciton (BSE_BND=T, bsebnds): the k+Q eigensolve, the device
assembly of matrizelbsekq (conduction bands at k+Q, exchange for |Q| != 0,
with and without temperature) and its eigenvalues against the oracle."""
import sys

import numpy as np
import pytest

from conftest import hex_model
from oracle import oracle as O
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu

COUL = dict(kind="V2DRK", r0=40.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001)
QFRAC = (0.125, -0.0625, 0.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(mesh, qfrac, nv=2, nc=2):
    m = hex_model(mesh)
    k = T.mp_mesh(*mesh, m.rlat)
    q = np.asarray(qfrac, float) @ T.recip_vectors(m.rlat)
    kq = T.mp_mesh_q(*mesh, m.rlat, q)
    ctx = hip.Context(0)
    ctx.set_tb_model(m)
    ctx.set_kpoints(k, mesh)
    ctx.hk_eig(nv, nc)
    bq = ctx.hk_eig_q(q, kq, nv, nc)
    om = O.Model(m, flags=hip.FLAG_COMPAT_F32 | hip.FLAG_GAUGE_LAPACK)
    ref, refq = O.hk_eig(om, k, nv, nc), O.hk_eig(om, kq, nv, nc)
    return m, k, q, ctx, ref, refq, bq


def oracle_matrix(m, k, q, ref, refq, mesh, temp=None, nv=2, nc=2):
    return O.bse_matrix_q(k, ref["eig"], ref["vec"], refq["eig"], refq["vec"], O.coulomb(**COUL), mesh, m.rlat,
                          nc, nv, q, temp=temp)


@pytest.mark.parametrize("temp", [None, 8000.0])
@pytest.mark.parametrize("precision", [hip.FP64, hip.MIXED])
def test_finite_q_matrix(precision, temp):
    say("case finite-Q matrix", precision, temp)
    mesh = (8, 8, 1)
    m, k, q, ctx, ref, refq, bq = setup(mesh, QFRAC)
    eq, _, nocpq = bq
    assert np.abs(eq - refq["eig"]).max() < 1e-10
    assert np.array_equal(nocpq, refq["nocp"])
    if temp is not None:
        ctx.set_temperature(temp)
    ctx.bse_build(COUL, precision)
    N = ctx.N
    H = ctx.get_block(0, N, 0, N)
    Hr = oracle_matrix(m, k, q, ref, refq, mesh, temp)
    Hx = O.bse_matrix_q(k, ref["eig"], ref["vec"], refq["eig"], refq["vec"], O.coulomb(**COUL), mesh, m.rlat,
                        2, 2, q, temp=temp, exchange=False)
    scale = np.abs(Hr).max()
    say("exchange size", np.abs(Hr - Hx).max(), "scale", scale)
    assert np.abs(Hr - Hx).max() > 1e-4          # the exchange term is visible here
    tol = 1e-12 * max(1.0, scale) if precision == hip.FP64 else 2e-6 * scale
    err = np.abs(H - Hr)
    np.fill_diagonal(err, 0.0)
    assert err.max() < tol
    assert np.array_equal(H, H.conj().T)
    assert np.abs(ctx.diag() - np.diag(Hr).real).max() < 1e-12 * max(1.0, scale)
    ctx.close()


@pytest.mark.parametrize("precision", [hip.FP64, hip.MIXED])
def test_finite_q_at_zero_is_plain_bse(precision):
    """Q = 0: the k+Q mesh is the k mesh and there is no exchange, so the
    store equals the plain one bit for bit (matrizelbsekq's modq = 0 branch)."""
    say("case finite-Q zero", precision)
    mesh = (6, 6, 1)
    m, k, q, ctx, ref, refq, bq = setup(mesh, (0.0, 0.0, 0.0))
    ctx.bse_build(COUL, precision)
    H = ctx.get_block(0, ctx.N, 0, ctx.N)
    d = ctx.diag()
    ctx.hk_eig_q(None, None, 2, 2)
    ctx.bse_build(COUL, precision)
    assert np.array_equal(ctx.get_block(0, ctx.N, 0, ctx.N), H)
    assert np.array_equal(ctx.diag(), d)
    ctx.close()


@pytest.mark.parametrize("mesh,nev", [((8, 8, 1), 256), ((20, 20, 1), 16)])
def test_finite_q_solve(mesh, nev):
    """Every eigenvalue at N = 256 (the reference's full ZHEEV per Q) and the
    lowest 16 at N = 1600 (LOBPCG) against the oracle matrix."""
    say("case finite-Q solve", mesh, nev)
    m, k, q, ctx, ref, refq, bq = setup(mesh, QFRAC)
    ctx.bse_build(COUL, hip.FP64)
    E, X, rep = ctx.bse_solve(nev)
    Hr = oracle_matrix(m, k, q, ref, refq, mesh)
    w = np.linalg.eigvalsh(Hr)[:nev]
    say("max |dE|", np.abs(E - w).max(), rep)
    assert np.abs(E - w).max() < 1e-9
    r = np.linalg.norm(Hr @ X - X * E, axis=0) / np.abs(E)
    assert r.max() < 1e-8
    ctx.close()


def test_finite_q_state_rules():
    say("case finite-Q state")
    mesh = (6, 6, 1)
    m, k, q, ctx, ref, refq, bq = setup(mesh, QFRAC)
    ctx.bse_build(COUL, hip.FP64)
    ctx.bse_solve(4)
    ctx.optics(0.05)
    with pytest.raises(hip.WTBError) as e:
        ctx.oscillator(4, "x", "x")
    assert e.value.code == -7
    with pytest.raises(hip.WTBError) as e:
        ctx.grad_energies(np.ones(4))
    assert e.value.code == -7
    # bands of another selection on the k mesh: the k+Q set no longer matches
    ctx.hk_eig(1, 1)
    with pytest.raises(hip.WTBError) as e:
        ctx.bse_build(COUL, hip.FP64)
    assert e.value.code == -6
    # new k-points switch finite Q off: a plain build again
    ctx.set_kpoints(k, mesh)
    ctx.hk_eig(2, 2)
    ctx.bse_build(COUL, hip.FP64)
    K = O.bse_matrix(k, ref["eig"], ref["vec"], O.coulomb(**COUL), mesh, m.rlat, 2, 2)
    assert np.abs(ctx.get_block(0, ctx.N, 0, ctx.N) - K).max() < 1e-12 * max(1.0, np.abs(K).max())
    ctx.close()
