"""The benchmark configurations through the C-ABI, anchored to the oracle.

C1  honeycomb-24: two-band honeycomb, 24x24 mesh, nv = nc = 1, N = 576, full
    diagonalisation (the reference's mkl_solver semantics), and the driver's
    1000-point spectrum over 3-7 eV against the oracle's optdiel/imagrenorm.
C2  hex22-60: 22-orbital hexagonal model, 60x60 mesh, nv = nc = 2, N = 14400,
    FP64, lowest 64 pairs and their oscillator strengths.
C3  hex22-90: the same model on 90x90, N = 32400, mixed precision against
    FP64 (0.1 meV) and both against the oracle.

Every solve is checked three ways: the energies against the committed LAPACK
fixtures of tests/golden (oracle FP64 matrix, zheevr; make_bse_golden.py),
the residual ||H_oracle x - E x|| of every returned vector through the
oracle's own FP64 matrix (all rows, built on the host), and, where both run,
mixed against FP64.  A Hermitian matrix has an eigenvalue within ||r|| of
every Ritz value, so the residual check bounds the energies against the
oracle matrix independently of the fixtures.
"""
import os
import sys
import time

import numpy as np
import pytest

from golden.make_bse_golden import golden_path, load as load_golden, model_fingerprint
from oracle import oracle as O
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "wtb_amd", "wtb_bse")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def prepare(name):
    model, spec = T.workload(name)
    g = load_golden(name)
    assert g["model_sha256"] == model_fingerprint(model), "fixture made from another model"
    mesh = tuple(spec["mesh"])
    k = T.mp_mesh(*mesh, model.rlat)
    return model, spec, g, mesh, k


def oracle_rows(model, spec, mesh, k):
    """the oracle's full FP64 BSE matrix (row-major) and its dipoles"""
    om = O.Model(model)
    r = O.hk_eig(om, k, spec["nv"], spec["nc"])
    assert r["status"] == 0
    N = k.shape[0] * spec["nv"] * spec["nc"]
    t = time.perf_counter()
    h = O.bse_rows(k, r["eig"], r["vec"], O.coulomb(**spec["coulomb"]), mesh, model.rlat, spec["nc"],
                   spec["nv"], 0, N)
    say(f"oracle matrix N={N} in {time.perf_counter() - t:.1f} s")
    return h


def rel_residuals(h, X, E):
    R = O.matvec_rows(h, X) - X * E[None, :]
    return np.linalg.norm(R, axis=0) / np.abs(E)


def level_sums(E, f, tol=1e-6):
    """oscillator strengths summed over each (near-)degenerate level"""
    cut = np.where(np.diff(E) > tol)[0] + 1
    return np.array([s.sum() for s in np.split(f, cut)]), cut


def test_c1_honeycomb24_full_diagonalisation():
    say("case C1 honeycomb-24 dense")
    model, spec, g, mesh, k = prepare("honeycomb-24")
    ctx = hip.Context(0)
    ctx.set_tb_model(model)
    ctx.set_kpoints(k, mesh)
    ctx.hk_eig(1, 1)
    ctx.bse_build(spec["coulomb"], hip.FP64)
    N = ctx.N
    assert N == 576
    E, X, rep = ctx.bse_solve(N)
    w = np.array(g["eigenvalues_eV"])
    say("max |E - LAPACK fixture| eV", np.abs(E - w).max())
    assert np.abs(E - w).max() < 1e-9
    h = oracle_rows(model, spec, mesh, k)
    assert np.abs(np.linalg.eigvalsh(h) - E).max() < 1e-9
    r = rel_residuals(h, X, E)
    say("max rel residual vs oracle", r.max())
    assert r.max() < 1e-10
    assert np.allclose(X.conj().T @ X, np.eye(N), atol=1e-10)
    # oscillator strengths, summed over degenerate levels (the honeycomb
    # mesh is symmetric: the eigenvectors inside a level are not unique)
    ctx.optics(spec["cshift"])
    for comp, key in (("x", "osc_x"), ("y", "osc_y")):
        f = ctx.oscillator(N, comp)
        got, cut = level_sums(E, f)
        ref, cut_r = level_sums(w, np.array(g[key]))
        assert np.array_equal(cut, cut_r)
        assert np.abs(got - ref).max() < 1e-8 * max(1.0, np.abs(ref).max())
    ctx.close()


def test_c1_honeycomb24_driver_spectrum():
    """The driver on the C1 cards: full diagonalisation, 1000-point spectrum
    over 3-7 eV with 0.05 eV broadening, checked against the oracle spectrum
    of the driver's own 4-decimal bse_opt_diel.dat and the fixture energies."""
    say("case C1 honeycomb-24 driver")
    if not os.path.exists(DRIVER):
        pytest.fail("wtb_amd/wtb_bse is not built (make -C wtb_amd/fortran)")
    from test_gpu_driver import load, run_driver
    g = load_golden("honeycomb-24")
    cards = ["BSE= T", "SPEC= T", "DTDIAG= T", "SYSDIM= \"2D\"", "NGX= 24", "NGY= 24", "NGZ= 1",
             "NBANDSC= 1", "NBANDSV= 1", "COULOMB_POT= V2DRK", "R_0= 10.0", "EDIEL= 1.0", "EDIEL_T= 1.0",
             "EDIEL_B= 1.0", "EDIEL_Z= 1.0", "LC= 1.0", "KTOL= 0.001", "CSHIFT= 0.05", "ENSPECI= 3.0",
             "ENSPECF= 7.0", "NESPEC= 1000", "NEV= 576", "PRECISION= FP64"]
    d, m = run_driver(cards)
    assert model_fingerprint(m) == g["model_sha256"]
    out = load(os.path.join(d, "bse_opt_diel.dat"))
    assert out.shape == (576, 7)
    w = np.array(g["eigenvalues_eV"])
    assert np.abs(out[:, 0] - w).max() <= 5.1e-5                 # F15.4 rounding
    vc = abs(np.dot(np.cross(m.rlat[0], m.rlat[1]), m.rlat[2]))
    for name, col, diag in (("xx", 1, True), ("yy", 2, True), ("xy", 4, False)):
        got = load(os.path.join(d, f"bse_diel_{name}.dat"))
        exp = O.spectrum(out[:, 0], out[:, col], diag, (24, 24, 1), vc, 3.0, 7.0, 1000, 0.05, renorm=True)
        assert got.shape == (1000, 3)
        assert np.abs(got[:, 0] - exp[:, 0]).max() < 1e-12
        scale = max(1.0, 2.0 * np.abs(exp[:, 1:]).max())
        assert np.abs(got[:, 1] - np.float32(2.0 * exp[:, 1])).max() < 1e-6 * scale, name
        assert np.abs(got[:, 2] - np.float32(2.0 * exp[:, 2])).max() < 1e-6 * scale, name
    absorb = load(os.path.join(d, "bse_absorption_coef.dat"))
    assert absorb.shape == (1000, 7) and np.all(np.isfinite(absorb[:, 1]))


def _solve_checked(ctx, prec, nev, tol):
    t = time.perf_counter()
    ctx.bse_build(COUL_CACHE["coulomb"], prec)
    E, X, rep = ctx.bse_solve(nev, tol=tol)
    say(f"precision {prec}: build + solve {time.perf_counter() - t:.3f} s, iterations {rep['iterations']}, "
        f"columns {rep['matvec_columns']}, max rel residual {rep['max_rel_residual']:.2e}")
    assert rep["nconv"] == nev
    assert rep["max_rel_residual"] < tol
    return E, X, rep


COUL_CACHE = {}


def test_c2_hex22_60_fp64_pairs_and_oscillators():
    say("case C2 hex22-60 FP64")
    model, spec, g, mesh, k = prepare("hex22-60")
    nev = 64
    COUL_CACHE["coulomb"] = spec["coulomb"]
    ctx = hip.Context(0)
    ctx.set_tb_model(model)
    ctx.set_kpoints(k, mesh)
    ctx.hk_eig(2, 2, host=False)
    assert ctx.N == 14400
    E, X, rep = _solve_checked(ctx, hip.FP64, nev, 1e-8)
    w = np.array(g["eigenvalues_eV"])
    err = np.abs(E - w).max()
    say("max |E - LAPACK fixture| eV", err)
    assert err < 1e-8
    # oscillator strengths: nondegenerate states (spacing >= 1 meV here), so
    # per state; the vectors carry errors ~ residual |E| / spacing
    ctx.optics(spec["cshift"])
    for comp, key in (("x", "osc_x"), ("y", "osc_y")):
        f = ctx.oscillator(nev, comp)[:nev]
        ref = np.array(g[key])
        say(comp, "max rel oscillator error", (np.abs(f - ref) / np.abs(ref).max()).max())
        assert np.abs(f - ref).max() < 1e-4 * np.abs(ref).max()
    h = oracle_rows(model, spec, mesh, k)
    r = rel_residuals(h, X, E)
    say("max rel residual vs oracle", r.max())
    assert r.max() < 1e-8
    # mixed on the same workload: 0.1 meV against the fixture
    Em, Xm, repm = _solve_checked(ctx, hip.MIXED, nev, 1e-5)
    say("mixed max |E - fixture| meV", np.abs(Em - w).max() * 1e3)
    assert np.abs(Em - w).max() < 1e-4
    rm = rel_residuals(h, Xm, Em)
    say("mixed max rel residual vs oracle", rm.max())
    assert rm.max() < 2e-5
    ctx.close()


def test_c3_hex22_90_mixed_vs_fp64():
    say("case C3 hex22-90 mixed vs FP64")
    if not os.path.exists(golden_path("hex22-90")):
        pytest.fail("tests/golden/bse_hex22-90.json is missing (make_bse_golden.py hex22-90)")
    model, spec, g, mesh, k = prepare("hex22-90")
    nev = 64
    COUL_CACHE["coulomb"] = spec["coulomb"]
    ctx = hip.Context(0)
    ctx.set_tb_model(model)
    ctx.set_kpoints(k, mesh)
    ctx.hk_eig(2, 2, host=False)
    assert ctx.N == 32400
    Em, Xm, repm = _solve_checked(ctx, hip.MIXED, nev, 1e-5)
    E, X, rep = _solve_checked(ctx, hip.FP64, nev, 1e-8)
    w = np.array(g["eigenvalues_eV"])
    say("FP64 max |E - LAPACK fixture| eV", np.abs(E - w).max())
    say("mixed max |E - FP64| meV", np.abs(Em - E).max() * 1e3)
    assert np.abs(E - w).max() < 1e-8
    assert np.abs(Em - E).max() < 1e-4
    assert np.abs(Em - w).max() < 1e-4
    h = oracle_rows(model, spec, mesh, k)
    r = rel_residuals(h, X, E)
    rm = rel_residuals(h, Xm, Em)
    say("max rel residual vs oracle: FP64", r.max(), "mixed", rm.max())
    assert r.max() < 1e-8
    assert rm.max() < 2e-5
    del h
    ctx.close()
