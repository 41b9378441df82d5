"""Temperature BSE (BSET=T, bsesolvertemp): the device assembly, solve and
oscillator strengths against the oracle's restatement (matrizelbsetemp with
the factor f = f_v - f_c of the row state on j >= i, dielbseptemp)."""
import sys

import numpy as np
import pytest

from conftest import hex_model
from oracle import oracle as O
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu

COUL = dict(kind="V2DRK", r0=40.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001)
TEMP = 8000.0        # kB T = 0.69 eV: factors down to ~0.6 on the hex22 bands


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(mesh, nv=2, nc=2):
    m = hex_model(mesh)
    k = T.mp_mesh(*mesh, m.rlat)
    ctx = hip.Context(0)
    ctx.set_tb_model(m)
    ctx.set_kpoints(k, mesh)
    ctx.hk_eig(nv, nc)
    ref = O.hk_eig(O.Model(m, flags=hip.FLAG_COMPAT_F32 | hip.FLAG_GAUGE_LAPACK), k, nv, nc, full=True)
    return m, k, ctx, ref


@pytest.mark.parametrize("precision", [hip.FP64, hip.MIXED])
def test_temperature_matrix(precision):
    say("case temperature matrix", precision)
    mesh = (8, 8, 1)
    m, k, ctx, ref = setup(mesh)
    ctx.set_temperature(TEMP)
    ctx.bse_build(COUL, precision)
    H = ctx.get_block(0, ctx.N, 0, ctx.N)
    Hr = O.bse_matrix_temp(k, ref["eig"], ref["vec"], O.coulomb(**COUL), mesh, m.rlat, 2, 2, TEMP)
    f, _ = O.state_factors(ref["eig"], 2, 2, TEMP)
    say("factor range", f.min(), f.max())
    assert f.min() < 0.9
    scale = np.abs(Hr).max()
    tol = 1e-12 * max(1.0, scale) if precision == hip.FP64 else 2e-6 * scale
    assert np.abs(H - Hr).max() < tol
    assert np.array_equal(H, H.conj().T)
    assert np.allclose(ctx.diag(), np.diag(Hr).real, rtol=0, atol=1e-12)
    ctx.close()


@pytest.mark.parametrize("precision", [hip.FP64, hip.MIXED])
def test_temperature_zero_is_plain_bse(precision):
    """T = 0 (1e-8 K in the reference) gives f = 1 at every state of this
    insulating model: the store equals the plain one bit for bit."""
    say("case temperature zero", precision)
    mesh = (6, 6, 1)
    m, k, ctx, ref = setup(mesh)
    ctx.bse_build(COUL, precision)
    H0 = ctx.get_block(0, ctx.N, 0, ctx.N)
    d0 = ctx.diag()
    ctx.set_temperature(0.0)
    ctx.bse_build(COUL, precision)
    assert np.array_equal(ctx.get_block(0, ctx.N, 0, ctx.N), H0)
    assert np.array_equal(ctx.diag(), d0)
    ctx.close()


@pytest.mark.parametrize("mesh,nev", [((8, 8, 1), 12), ((20, 20, 1), 16)])
def test_temperature_solve_and_oscillator(mesh, nev):
    """Dense path (N = 256) and LOBPCG (N = 1600): the lowest pairs of the
    temperature matrix and the f-weighted oscillator strengths."""
    say("case temperature solve", mesh, nev)
    m, k, ctx, ref = setup(mesh)
    ctx.set_temperature(TEMP)
    ctx.bse_build(COUL, hip.FP64)
    E, X, rep = ctx.bse_solve(nev)
    Hr = O.bse_matrix_temp(k, ref["eig"], ref["vec"], O.coulomb(**COUL), mesh, m.rlat, 2, 2, TEMP)
    w = np.linalg.eigvalsh(Hr)[:nev]
    say("max |dE|", np.abs(E - w).max(), rep)
    assert np.abs(E - w).max() < 1e-9
    # the solver's own criterion: relative residual norm per pair below tol = 1e-8
    r = np.linalg.norm(Hr @ X - X * E, axis=0) / np.abs(E)
    assert r.max() < 1e-8
    d = ctx.optics(0.05)
    f, _ = O.state_factors(ref["eig"], 2, 2, TEMP)
    for a, b in (("x", "x"), ("y", "y"), ("x", "y")):
        ia, ib = hip.COMPONENTS[a], hip.COMPONENTS[b]
        got = ctx.oscillator(nev, a, b)[:nev]
        want = np.real((X.T @ d[:, ia]) * np.conj(X.T @ (d[:, ib] * f)))
        assert np.abs(got - want).max() <= 1e-10 * max(1e-30, np.abs(want).max())
    ctx.close()


def test_temperature_switch_and_adjoint():
    say("case temperature switch")
    mesh = (6, 6, 1)
    m, k, ctx, ref = setup(mesh)
    ctx.set_temperature(TEMP)
    ctx.bse_build(COUL, hip.FP64)
    ctx.bse_solve(4)
    with pytest.raises(hip.WTBError) as e:
        ctx.grad_energies(np.ones(4))
    assert e.value.code == -7
    # another temperature drops the store: it must be built again
    ctx.set_temperature(300.0)
    with pytest.raises(hip.WTBError) as e:
        ctx.bse_solve(4)
    assert e.value.code == -6
    ctx.set_temperature(None)
    ctx.bse_build(COUL, hip.FP64)
    ctx.bse_solve(4)
    ctx.close()
