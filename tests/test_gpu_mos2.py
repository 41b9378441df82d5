"""The reference's own example input on the HIP path: examples/example-3
(tb_mos2.dat, MoS2 with spin-orbit coupling: norb = 22, 147 R vectors, six
of them with Wigner-Seitz degeneracy 2, so H(k) = sum_R e^{ik.R} H(R)/deg
runs with deg != 1).

* H(k) bands, the BSE matrix (V2DT2, LC = 8) and the dipoles on a 6x6 mesh
  against the oracle.
* example_3.sh's cards through the Fortran driver (RK2D, RK = 120 -> 44x44,
  nv = nc = 2, N = 7744, 100 pairs as the reference's SLEPc call): the
  exciton energies of bse_opt_diel.dat against the LAPACK fixture of the
  oracle matrix (tests/golden/bse_mos2-44.json) and the spectrum files
  against the oracle spectrum of the written energies.
* The same solve through the C-ABI, FP64 and mixed, with the residual of
  every vector through the oracle's FP64 matrix; the wall times are printed
  next to the reference's published 3 s (SLEPc) and 680 s (MKL ZHEEV) on 64
  cores (README.md:20).
"""
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np
import pytest

from golden.make_bse_golden import MOS2_FILE, MOS2_SPEC, load as load_golden, model_fingerprint
from oracle import oracle as O
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "wtb_amd", "wtb_bse")
COUL = MOS2_SPEC["coulomb"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_up_to_sign(a, b, tol):
    ov = np.sum(a.conj() * b, axis=-1)
    s = np.where(ov.real < 0, -1.0, 1.0)
    return np.abs(a * s[..., None] - b).max() < tol


def test_mos2_bands_kernel_and_dipoles_small_mesh():
    say("case MoS2 6x6 bands, kernel, dipoles")
    m = T.read_tb_file(MOS2_FILE)
    assert np.count_nonzero(m.deg == 2) == 6
    mesh = (6, 6, 1)
    k = T.mp_mesh(*mesh, m.rlat)
    flags = hip.FLAG_COMPAT_F32 | hip.FLAG_GAUGE_LAPACK
    ctx = hip.Context(0)
    ctx.set_tb_model(m, flags)
    ctx.set_kpoints(k, mesh)
    eig, vec, nocp = ctx.hk_eig(2, 2)
    om = O.Model(m, flags=flags)
    ref = O.hk_eig(om, k, 2, 2, full=True)
    assert np.array_equal(nocp, ref["nocp"]) and np.all(nocp == 14)
    assert np.abs(eig - ref["eig"]).max() < 1e-10
    e_all, _ = ctx.get_bands()
    assert np.abs(e_all - ref["eall"]).max() < 1e-10
    assert same_up_to_sign(vec, ref["vec"], 1e-8)
    # The two valence and the two conduction bands are spin-orbit pairs split
    # by only ~6e-6 eV on this mesh, so each band vector is fixed to about
    # eps |H| / spacing and the kernel elements (products of overlaps of such
    # vectors) agree to that level, not to 1e-12: the FP64 tolerance is scaled
    # by 1e-4 eV / spacing.  The eigenvalues of the kernel do not depend on
    # the vectors' freedom inside a near-degenerate pair and are compared too.
    sel = ref["eall"][:, 12:16]
    spacing = min(np.diff(sel[:, :2], axis=1).min(), np.diff(sel[:, 2:], axis=1).min())
    say("smallest spin-orbit spacing eV", spacing)
    Hr = O.bse_matrix(k, ref["eig"], ref["vec"], O.coulomb(**COUL), mesh, m.rlat, 2, 2)
    wr = np.linalg.eigvalsh(Hr)
    for prec in (hip.FP64, hip.MIXED):
        ctx.bse_build(COUL, prec)
        H = ctx.get_block(0, ctx.N, 0, ctx.N)
        scale = np.abs(Hr).max()
        tol = 1e-12 * max(1.0, scale) * max(1.0, 1e-4 / spacing) if prec == hip.FP64 else 2e-6 * scale
        say("precision", prec, "max |H - H_oracle|", np.abs(H - Hr).max(), "tol", tol)
        assert np.abs(H - Hr).max() < tol
        assert np.array_equal(H, H.conj().T)
        dw = np.abs(np.linalg.eigvalsh(H) - wr).max()
        assert dw < (1e-11 if prec == hip.FP64 else 1e-4), dw
    d = ctx.optics(MOS2_SPEC["cshift"])
    dr = O.dipoles(om, k, ref["eig"], ref["vec"], 2, 2, MOS2_SPEC["cshift"])
    assert np.abs(np.abs(d.T) - np.abs(dr)).max() < 1e-9
    ctx.close()


def test_mos2_example3_driver_run():
    say("case MoS2 example-3 driver (RK2D 120, N = 7744)")
    if not os.path.exists(DRIVER):
        pytest.fail("wtb_amd/wtb_bse is not built (make -C wtb_amd/fortran)")
    g = load_golden("mos2-44")
    d = tempfile.mkdtemp(prefix="wtbex3")
    shutil.copy(MOS2_FILE, os.path.join(d, "tb_mos2.dat"))
    # examples/example-3/example_3.sh, with OUTPUT in the run folder
    cards = ["NTHREADS= 10", 'SYSDIM= "2D"', 'DFT= "V"', f'OUTPUT= "{d}/"', f'CALC_DATA= "{d}/"',
             'PARAMS_FILE= "tb_mos2.dat"', 'MESH_TYPE= "RK2D"', "RK= 120", "BSE= T", "SPEC= T", "DTDIAG= T",
             "COULOMB_POT= V2DT2", "NBANDSC= 2", "NBANDSV= 2", "LC= 8.00", "CSHIFT= 0.08", "ENSPECI= 0.0",
             "ENSPECF= 4.0"]
    t = time.perf_counter()
    r = subprocess.run([DRIVER], input="\n".join(cards) + "\n", capture_output=True, text=True, timeout=300, cwd=d)
    wall = time.perf_counter() - t
    say(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    say(f"example-3 driver wall time {wall:.2f} s (reference: 3 s SLEPc, 680 s MKL on 64 cores)")
    out = np.loadtxt(os.path.join(d, "bse_opt_diel.dat"), comments="#", ndmin=2)
    assert out.shape == (7744, 7)
    w = np.array(g["eigenvalues_eV"])
    assert np.abs(out[:100, 0] - w).max() <= 5.1e-5          # F15.4 rounding
    assert np.all(out[100:, :] == 0.0)
    m = T.read_tb_file(MOS2_FILE)
    vc = abs(np.dot(np.cross(m.rlat[0], m.rlat[1]), m.rlat[2]))
    got = np.loadtxt(os.path.join(d, "bse_diel_xx.dat"), ndmin=2)
    exp = O.spectrum(out[:, 0], out[:, 1], True, (44, 44, 1), vc, 0.0, 4.0, got.shape[0], 0.08, renorm=True)
    assert np.abs(got[:, 0] - exp[:, 0]).max() < 1e-12
    scale = max(1.0, np.abs(exp[:, 1:]).max())
    assert np.abs(got[:, 1] - np.float32(exp[:, 1])).max() < 1e-6 * scale      # SOC: no spin factor
    assert np.abs(got[:, 2] - np.float32(exp[:, 2])).max() < 1e-6 * scale


def test_mos2_example3_solve_fp64_and_mixed():
    say("case MoS2 example-3 solve through the C-ABI")
    g = load_golden("mos2-44")
    m = T.read_tb_file(MOS2_FILE)
    assert g["model_sha256"] == model_fingerprint(m)
    mesh = (44, 44, 1)
    k = T.mp_mesh(*mesh, m.rlat)
    ctx = hip.Context(0)
    ctx.set_tb_model(m)
    ctx.set_kpoints(k, mesh)
    nev = 100
    res = {}
    for prec, tol, name in ((hip.FP64, 1e-8, "fp64"), (hip.MIXED, 1e-5, "mixed"), (hip.FP64, 1e-8, "fp64"),
                            (hip.MIXED, 1e-5, "mixed")):
        t = time.perf_counter()
        ctx.hk_eig(2, 2, host=False)
        ctx.bse_build(COUL, prec)
        E, X, rep = ctx.bse_solve(nev, tol=tol)
        dt = time.perf_counter() - t
        say(f"{name}: H(k) + build + solve {dt * 1e3:.1f} ms (second run is warm), iterations {rep['iterations']}, "
            f"max rel residual {rep['max_rel_residual']:.2e}")
        assert rep["nconv"] == nev and rep["max_rel_residual"] < tol
        res[name] = (E, X)
    w = np.array(g["eigenvalues_eV"])
    E64, X64 = res["fp64"]
    Em, Xm = res["mixed"]
    say("FP64 max |E - fixture| eV", np.abs(E64 - w).max(), "mixed max |E - FP64| meV", np.abs(Em - E64).max() * 1e3)
    assert np.abs(E64 - w).max() < 1e-8
    assert np.abs(Em - E64).max() < 1e-4
    om = O.Model(m)
    r = O.hk_eig(om, k, 2, 2)
    h = O.bse_rows(k, r["eig"], r["vec"], O.coulomb(**COUL), mesh, m.rlat, 2, 2, 0, 7744)
    for name, (E, X), lim in (("fp64", res["fp64"], 1e-8), ("mixed", res["mixed"], 2e-5)):
        rr = np.linalg.norm(O.matvec_rows(h, X) - X * E[None, :], axis=0) / np.abs(E)
        say(name, "max rel residual vs oracle", rr.max())
        assert rr.max() < lim
    ctx.close()
