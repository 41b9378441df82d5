"""The Fortran driver (wtb_amd/wtb_bse, ISO_C_BINDING on the HIP library) on the
honeycomb case: its output files against the oracle.

bse_opt_diel.dat: exciton energies against the eigenvalues of the oracle
matrix and the oscillator strengths summed over each degenerate level (the
sum is invariant under rotations inside a degenerate eigenspace);
sp_optics.dat: the sorted transition energies and |d|^2;
bse_diel_*.dat: the oracle spectrum (optdiel + imagrenorm) evaluated on the
driver's own 4-decimal bse_opt_diel.dat, as the reference re-reads it;
bseoptprop files and exciton_lifetime.dat: the closed-form optical constants
and lifetimes of those spectra.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "wtb_amd", "wtb_bse")
PI32 = float(np.float32(np.arccos(np.float32(-1.0))))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(path):
    return np.loadtxt(path, comments="#", ndmin=2)


def run_driver(cards):
    d = tempfile.mkdtemp(prefix="wtb")
    m = T.honeycomb_two_band()
    T.write_tb_file(m, os.path.join(d, "tb.dat"))
    text = "\n".join([f'PARAMS_FILE= "{d}/tb.dat"', f'OUTPUT= "{d}/"'] + cards) + "\n"
    say("driver cards:", text.replace("\n", " | "))
    r = subprocess.run([DRIVER], input=text, capture_output=True, text=True, timeout=240, cwd=d)
    say(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    return d, m


@pytest.fixture(scope="module")
def honeycomb_run():
    if not os.path.exists(DRIVER):
        pytest.fail("wtb_amd/wtb_bse is not built (make -C wtb_amd/fortran)")
    cards = ["BSE= T", "SPEC= T", "DTDIAG= T", "CPOL= T", "SYSDIM= \"2D\"", "NGX= 12", "NGY= 12", "NGZ= 1",
             "NBANDSC= 1", "NBANDSV= 1", "COULOMB_POT= V2DRK", "R_0= 10.0", "EDIEL= 1.0", "EDIEL_T= 1.0",
             "EDIEL_B= 1.0", "EDIEL_Z= 1.0", "LC= 1.0", "KTOL= 0.001", "CSHIFT= 0.05", "ENSPECI= 3.0",
             "ENSPECF= 7.0", "NESPEC= 400", "NEV= 144", "PRECISION= FP64", "BSE_WF= T", "EXC_WF_I= 1",
             "EXC_WF_F= 4"]
    return run_driver(cards)


def oracle_case(m):
    mesh = (12, 12, 1)
    k = T.mp_mesh(*mesh, m.rlat)
    om = O.Model(m)
    ref = O.hk_eig(om, k, 1, 1)
    c = O.coulomb(kind="V2DRK", r0=10.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001)
    H = O.bse_matrix(k, ref["eig"], ref["vec"], c, mesh, m.rlat, 1, 1)
    w, X = np.linalg.eigh(H)
    d = O.dipoles(om, k, ref["eig"], ref["vec"], 1, 1, 0.05)      # (5, N): x, y, z, sp, sm
    return mesh, k, ref, w, X, d


def test_bse_opt_diel_energies_and_strengths(honeycomb_run):
    d, m = honeycomb_run
    say("case driver bse_opt_diel")
    mesh, k, ref, w, X, dip = oracle_case(m)
    out = load(os.path.join(d, "bse_opt_diel.dat"))
    assert out.shape == (144, 7)
    assert np.abs(out[:, 0] - w).max() <= 5.1e-5            # F15.4 rounding
    # oscillator strengths summed over each degenerate level
    levels = np.split(np.arange(144), np.where(np.diff(w) > 1e-6)[0] + 1)
    for comp, col in ((0, 1), (1, 2), (2, 3)):
        fr = np.abs(X.T @ dip[comp]) ** 2
        for lv in levels:
            assert abs(out[lv, col].sum() - fr[lv].sum()) <= 5.1e-5 * len(lv) + 1e-6 * fr[lv].sum()
    pol = load(os.path.join(d, "bse_opt_diel-pol.dat"))
    assert pol.shape == (144, 6)
    for comp, col in ((3, 4), (4, 5)):
        fr = np.abs(X.T @ dip[comp]) ** 2
        for lv in levels:
            assert abs(pol[lv, col].sum() - fr[lv].sum()) <= 5.1e-5 * len(lv) + 1e-6 * fr[lv].sum()


def test_sp_optics_sorted_transitions(honeycomb_run):
    d, m = honeycomb_run
    say("case driver sp_optics")
    mesh, k, ref, w, X, dip = oracle_case(m)
    sp = load(os.path.join(d, "sp_optics.dat"))
    de = (ref["eig"][:, 1] - ref["eig"][:, 0]).astype(np.float32)
    order = np.argsort(de, kind="stable")
    assert np.abs(sp[:, 0] - de[order]).max() <= 5.1e-5
    # transitions of equal energy may come in either order: compare the
    # (energy, |d_x|^2, |d_y|^2) rows as sorted sets, on the written 4 decimals
    ax = (np.abs(dip[0]) ** 2).astype(np.float32)
    ay = (np.abs(dip[1]) ** 2).astype(np.float32)
    mine = np.round(np.stack([de, ax, ay], 1).astype(np.float64), 4)
    mine = mine[np.lexsort(mine.T[::-1])]
    theirs = sp[:, :3][np.lexsort(sp[:, :3].T[::-1])]
    assert np.abs(mine - theirs).max() <= 1.01e-4


def test_spectra_match_oracle_on_the_written_file(honeycomb_run):
    d, m = honeycomb_run
    say("case driver spectra")
    out = load(os.path.join(d, "bse_opt_diel.dat"))
    vc = abs(np.dot(np.cross(m.rlat[0], m.rlat[1]), m.rlat[2]))
    spinf = 2.0                                   # NP model
    for name, col, diag in (("xx", 1, True), ("yy", 2, True), ("zz", 3, True),
                            ("xy", 4, False), ("xz", 5, False), ("yz", 6, False)):
        got = load(os.path.join(d, f"bse_diel_{name}.dat"))
        exp = O.spectrum(out[:, 0], out[:, col], diag, (12, 12, 1), vc, 3.0, 7.0, 400, 0.05, renorm=True)
        assert got.shape == (400, 3)
        assert np.abs(got[:, 0] - exp[:, 0]).max() < 1e-12
        scale = max(1.0, np.abs(exp[:, 1:]).max() * spinf)
        # the driver prints float32 values list-directed
        assert np.abs(got[:, 1] - np.float32(spinf * exp[:, 1])).max() < 1e-6 * scale, name
        assert np.abs(got[:, 2] - np.float32(spinf * exp[:, 2])).max() < 1e-6 * scale, name


@pytest.mark.filterwarnings("ignore::RuntimeWarning")      # 0/0 on the identically zero xy, xz, yz
def test_optical_constants_and_lifetimes(honeycomb_run):
    d, m = honeycomb_run
    say("case driver optical constants")
    e = {c: load(os.path.join(d, f"bse_diel_{c}.dat")) for c in ("xx", "yy", "zz", "xy", "xz", "yz")}
    comps = ("xx", "yy", "zz", "xy", "xz", "yz")
    e1 = np.stack([e[c][:, 1] for c in comps], 1)
    e2 = np.stack([e[c][:, 2] for c in comps], 1)
    en = e["xx"][:, 0]
    mod = np.sqrt(e1 * e1 + e2 * e2)
    n = np.sqrt((mod + e1) / 2)
    kap = np.sqrt((mod - e1) / 2)
    hc = float(np.float32(19.746e-6))
    expect = {
        "bse_refractive_index.dat": n,
        "bse_extinction_coef.dat": kap,
        "bse_reflectibility.dat": ((n - 1) ** 2 + kap ** 2) / ((n + 1) ** 2 + kap ** 2),
        "bse_absorption_coef.dat": (float(np.sqrt(np.float32(2.0))) * en)[:, None] * (1.0 / hc) * np.sqrt(mod - e1),
        "bse_en_loss_func.dat": e2 / (e1 * e1 + e2 * e2),
    }
    for fname, val in expect.items():
        got = load(os.path.join(d, fname))
        assert got.shape == (400, 7)
        ok = np.isfinite(val)
        assert np.abs(got[:, 0] - en).max() <= 5.1e-5
        assert np.abs(got[:, 1:][ok] - val[ok]).max() <= 5.1e-5 + 1e-7 * np.abs(val[ok]).max(), fname
    # 2-D radiative lifetimes from the written exciton energies / strengths
    out = load(os.path.join(d, "bse_opt_diel.dat"))
    lt = load(os.path.join(d, "exciton_lifetime.dat"))
    cfe, hbar = float(np.float32(7.297e-3)), float(np.float32(6.582e-16))
    a2 = np.linalg.norm(np.cross(m.rlat[0], m.rlat[1]))
    f = out[:, 1]
    sel = f > 1e-3
    ref = (144 * a2 * hbar) / ((8.0 * PI32 * cfe * out[:, 0]) * f)
    assert np.allclose(lt[sel, 2], ref[sel], rtol=1e-3)


def test_temperature_driver_run():
    """BSET=T with the Varshni gap correction (TA= VE): the log names the
    temperature, exciton energies are the eigenvalues of the temperature
    matrix built on bands shifted by the correction, the strengths and
    sp_optics columns carry f_v - f_c."""
    say("case driver temperature")
    if not os.path.exists(DRIVER):
        pytest.fail("wtb_amd/wtb_bse is not built (make -C wtb_amd/fortran)")
    temp, st, ph = 12000.0, 5e-4, 0.02
    cards = ["BSET= T", f"TEMP= {temp}", 'TA= "VE"', f"ST= {st}", f"PHAVG= {ph}", "SPEC= T",
             "DTDIAG= T", "SYSDIM= \"2D\"", "NGX= 12", "NGY= 12", "NGZ= 1", "NBANDSC= 1", "NBANDSV= 1",
             "COULOMB_POT= V2DRK", "R_0= 10.0", "EDIEL= 1.0", "EDIEL_T= 1.0", "EDIEL_B= 1.0", "EDIEL_Z= 1.0",
             "LC= 1.0", "KTOL= 0.001", "CSHIFT= 0.05", "ENSPECI= 3.0", "ENSPECF= 7.0", "NESPEC= 400",
             "NEV= 144", "PRECISION= FP64"]
    d, m = run_driver(cards)
    with open(os.path.join(d, "log_bse-diel.dat")) as fh:
        first = fh.readline().split()
    assert first[0] == "temperature" and float(first[1]) == temp
    # no lifetimes on the temperature path (main.F90:187-190)
    assert not os.path.exists(os.path.join(d, "exciton_lifetime.dat"))
    assert os.path.exists(os.path.join(d, "sp_diel_xx.dat"))
    tcor = T.temperature_gap_shift("VE", st, ph, temp)
    say("gap correction", tcor)
    assert -2e-3 < tcor < -5e-4
    m2 = m.copy()
    m2.scissor = m.scissor + tcor
    mesh = (12, 12, 1)
    k = T.mp_mesh(*mesh, m.rlat)
    om = O.Model(m2)
    ref = O.hk_eig(om, k, 1, 1)
    c = O.coulomb(kind="V2DRK", r0=10.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001)
    H = O.bse_matrix_temp(k, ref["eig"], ref["vec"], c, mesh, m.rlat, 1, 1, temp)
    w, X = np.linalg.eigh(H)
    f, _ = O.state_factors(ref["eig"], 1, 1, temp)
    assert f.min() < 0.95
    dip = O.dipoles(om, k, ref["eig"], ref["vec"], 1, 1, 0.05)
    out = load(os.path.join(d, "bse_opt_diel.dat"))
    assert out.shape == (144, 7)
    assert np.abs(out[:, 0] - w).max() <= 5.1e-5
    levels = np.split(np.arange(144), np.where(np.diff(w) > 1e-6)[0] + 1)
    for comp, col in ((0, 1), (1, 2), (2, 3)):
        fr = np.real((X.T @ dip[comp]) * np.conj(X.T @ (dip[comp] * f)))
        for lv in levels:
            assert abs(out[lv, col].sum() - fr[lv].sum()) <= 5.1e-5 * len(lv) + 1e-6 * abs(fr[lv].sum())
    sp = load(os.path.join(d, "sp_optics.dat"))
    de = (ref["eig"][:, 1] - ref["eig"][:, 0]).astype(np.float32)
    ax = (np.abs(dip[0]) ** 2 * f).astype(np.float32)
    ay = (np.abs(dip[1]) ** 2 * f).astype(np.float32)
    mine = np.round(np.stack([de, ax, ay], 1).astype(np.float64), 4)
    mine = mine[np.lexsort(mine.T[::-1])]
    theirs = sp[:, :3][np.lexsort(sp[:, :3].T[::-1])]
    assert np.abs(mine - theirs).max() <= 1.01e-4


def test_exciton_bands_driver_run():
    """BSE_BND=T along one Q segment (Gamma -> b1/2, three points) on the
    honeycomb 6x6 mesh: every eigenvalue of the finite-Q matrix (with the
    exchange for Q != 0) per Q in bands_bse.dat, the path coordinate and the
    segment ends of KLABELS-BSE.dat."""
    say("case driver exciton bands")
    if not os.path.exists(DRIVER):
        pytest.fail("wtb_amd/wtb_bse is not built (make -C wtb_amd/fortran)")
    kdir = tempfile.mkdtemp(prefix="wtbq")
    kfile = os.path.join(kdir, "kpath_bse.dat")
    ks = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0]])
    with open(kfile, "w") as fh:
        fh.write("2\n3\n" + "".join(f"{a} {b} {c}\n" for a, b, c in ks))
    cards = ["BSE_BND= T", f'KPATH_BSE= "{kfile}"', "SYSDIM= \"2D\"", "NGX= 6", "NGY= 6", "NGZ= 1",
             "NBANDSC= 1", "NBANDSV= 1", "COULOMB_POT= V2DRK", "R_0= 10.0", "EDIEL= 1.0", "EDIEL_T= 1.0",
             "EDIEL_B= 1.0", "EDIEL_Z= 1.0", "LC= 1.0", "KTOL= 0.001", "PRECISION= FP64", "BSE_WF= T",
             "EXC_WF_I= 1", "EXC_WF_F= 3"]
    d, m = run_driver(cards)
    assert not os.path.exists(os.path.join(d, "bse_opt_diel.dat"))      # BSE= is off
    lab = load(os.path.join(d, "KLABELS-BSE.dat"))[:, 0]
    assert np.allclose(lab, [0.0, 1.0], atol=1e-7)
    with open(os.path.join(d, "log_bse_kpath.dat")) as fh:
        log = fh.read()
    assert log.count("progress:") == 3
    out = load(os.path.join(d, "bands_bse.dat"))
    N, nq = 36, 3
    assert out.shape == (N * nq, 2)
    mesh = (6, 6, 1)
    b = T.recip_vectors(m.rlat)
    kc = [(s[0] * b[0] + s[1] * b[1]) + s[2] * b[2] for s in ks]
    k = T.mp_mesh(*mesh, m.rlat)
    om = O.Model(m)
    ref = O.hk_eig(om, k, 1, 1)
    c = O.coulomb(kind="V2DRK", r0=10.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001)
    moved = 0.0
    for iq in range(nq):
        q = kc[0] + (kc[1] - kc[0]) * float(iq) / 2.0
        kq = T.mp_mesh_q(*mesh, m.rlat, q)
        refq = O.hk_eig(om, kq, 1, 1)
        H = O.bse_matrix_q(k, ref["eig"], ref["vec"], refq["eig"], refq["vec"], c, mesh, m.rlat, 1, 1, q)
        w = np.linalg.eigvalsh(H)
        rows = out[iq::nq]
        assert np.abs(rows[:, 0] - 0.5 * iq).max() < 1e-6
        assert np.abs(rows[:, 1] - w).max() < 1e-9, iq
        # BSE_WF=T (excwfi2): exc_wf_<n>_<iq>.dat for n = 1..3, headed by the
        # Cartesian Q; |psi|^2 summed over each level inside 1..3 against the
        # oracle's projector on that level
        wv, Xq = np.linalg.eigh(H)
        levels = [lv for lv in np.split(np.arange(N), np.where(np.diff(wv) > 1e-6)[0] + 1) if lv[-1] < 3]
        amp = {}
        for n in range(1, 4):
            f = os.path.join(d, f"exc_wf_{n}_{iq + 1}.dat")
            with open(f) as fh:
                head = fh.read().splitlines()
            assert np.allclose([float(x) for x in head[0].split()[-3:]], np.float32(q), atol=1e-6)
            assert abs(float(head[2].split()[-1]) - wv[n - 1]) < 1e-8
            a = load(f)
            assert a.shape == (N, 8)
            assert np.abs(a[:, 1:4] - k).max() <= 5.1e-5
            assert np.all(a[:, 4] == 2) and np.all(a[:, 5] == 1)
            amp[n] = a[:, 6] + 1j * a[:, 7]
        for lv in levels:
            got = sum(np.abs(amp[i + 1]) ** 2 for i in lv)
            want = (np.abs(Xq[:, lv]) ** 2).sum(1)
            assert np.abs(got - want).max() < 1e-3 * want.max(), (iq, lv)
        H0 = O.bse_matrix_q(k, ref["eig"], ref["vec"], refq["eig"], refq["vec"], c, mesh, m.rlat, 1, 1, q,
                            exchange=False)
        moved = max(moved, np.abs(np.linalg.eigvalsh(H0) - w).max())
    assert moved > 1e-4                # the exchange shifts the Q != 0 levels


@pytest.mark.filterwarnings("ignore::RuntimeWarning")      # 0/0 on identically zero components
def test_single_particle_and_polarised_spectra(honeycomb_run):
    """SPEC=T with BSE: the single-particle chains (spdielraw, spdielrawpol on
    sp_optics*.dat) and, with CPOL=T, bsedielrawpol on bse_opt_diel-pol.dat,
    each against the oracle spectrum of the driver's own 4-decimal file; their
    optical constants and the polarised lifetimes from the closed forms."""
    d, m = honeycomb_run
    say("case driver single-particle and polarised spectra")
    vc = abs(np.dot(np.cross(m.rlat[0], m.rlat[1]), m.rlat[2]))
    spinf = 2.0
    sets = [("sp_optics.dat", "sp_diel_", ("xx", "yy", "zz", "xy", "xz", "yz"), 3),
            ("sp_optics-pol.dat", "sp_diel-pol_", ("x", "y", "z", "sp", "sm"), 5),
            ("bse_opt_diel-pol.dat", "bse_diel-pol_", ("x", "y", "z", "sp", "sm"), 5)]
    for src, prefix, tags, ndiag in sets:
        tr = load(os.path.join(d, src))
        for col, tag in enumerate(tags, start=1):
            got = load(os.path.join(d, f"{prefix}{tag}.dat"))
            exp = O.spectrum(tr[:, 0], tr[:, col], col <= ndiag, (12, 12, 1), vc, 3.0, 7.0, 400, 0.05, renorm=True)
            assert got.shape == (400, 3), (prefix, tag)
            scale = max(1.0, np.abs(exp[:, 1:]).max() * spinf)
            assert np.abs(got[:, 0] - exp[:, 0]).max() < 1e-12
            assert np.abs(got[:, 1] - np.float32(spinf * exp[:, 1])).max() < 1e-6 * scale, (prefix, tag)
            assert np.abs(got[:, 2] - np.float32(spinf * exp[:, 2])).max() < 1e-6 * scale, (prefix, tag)
    # optical constants of the single-particle and polarised sets
    for inp, outp, suffix, tags in (("sp_diel_", "sp_", "", ("xx", "yy", "zz", "xy", "xz", "yz")),
                                    ("sp_diel-pol_", "sp_", "-pol", ("x", "y", "z", "sp", "sm")),
                                    ("bse_diel-pol_", "bse_", "-pol", ("x", "y", "z", "sp", "sm"))):
        e = [load(os.path.join(d, f"{inp}{t}.dat")) for t in tags]
        e1 = np.stack([x[:, 1] for x in e], 1)
        e2 = np.stack([x[:, 2] for x in e], 1)
        mod = np.sqrt(e1 * e1 + e2 * e2)
        n = np.sqrt((mod + e1) / 2)
        got = load(os.path.join(d, f"{outp}refractive_index{suffix}.dat"))
        assert got.shape == (400, len(tags) + 1)
        assert np.abs(got[:, 1:] - n).max() <= 5.1e-5 + 1e-7 * np.abs(n).max(), (outp, suffix)
        loss = e2 / (e1 * e1 + e2 * e2)
        got = load(os.path.join(d, f"{outp}en_loss_func{suffix}.dat"))
        ok = np.isfinite(loss)
        assert np.abs(got[:, 1:][ok] - loss[ok]).max() <= 5.1e-5 + 1e-7 * np.abs(loss[ok]).max()
    # polarised lifetimes: the same closed form per column, the first on x + y + z
    pol = load(os.path.join(d, "bse_opt_diel-pol.dat"))
    lt = load(os.path.join(d, "exciton_lifetime-pol.dat"))
    assert lt.shape == (144, 7)
    cfe, hbar = float(np.float32(7.297e-3)), float(np.float32(6.582e-16))
    a2 = np.linalg.norm(np.cross(m.rlat[0], m.rlat[1]))
    for col, f in ((1, pol[:, 1] + pol[:, 2] + pol[:, 3]), (5, pol[:, 4]), (6, pol[:, 5])):
        sel = f > 1e-3
        ref = (144 * a2 * hbar) / ((8.0 * PI32 * cfe * pol[:, 0]) * f)
        assert np.allclose(lt[sel, col], ref[sel], rtol=1e-3), col


def test_dos_driver_run():
    """DOS=T alone (no BSE): dos.dat of the honeycomb model against the
    oracle density (spin factor 2 of an NP model) and its running trapezoid
    integral, written one row late as dostool writes it."""
    say("case driver dos")
    if not os.path.exists(DRIVER):
        pytest.fail("wtb_amd/wtb_bse is not built (make -C wtb_amd/fortran)")
    cards = ["DOS= T", "NEDOS= 401", "SIGMA= 0.1", "NGX= 12", "NGY= 12", "NGZ= 1"]
    d, m = run_driver(cards)
    assert not os.path.exists(os.path.join(d, "bse_opt_diel.dat"))
    out = load(os.path.join(d, "dos.dat"))
    assert out.shape == (401, 4)
    k = T.mp_mesh(12, 12, 1, m.rlat)
    r = O.hk_eig(O.Model(m), k, 1, 1, full=True)
    en, dd = O.dos(r["eall"], r["zall"], 0.1, 401)
    assert np.abs(out[:, 0] - en).max() <= 5.1e-5
    assert np.abs(out[:, 1] - 2.0 * dd[:, 0]).max() <= 5.1e-5
    assert np.abs(out[:, 3] - 2.0 * dd[:, 3]).max() <= 5.1e-5
    tdos = 2.0 * dd[:, 0]
    run = np.r_[0.0, 0.0, np.cumsum(0.5 * (tdos[1:-1] + tdos[:-2]) * (en[-1] - en[0]) / 400.0)]
    assert np.abs(out[:, 2] - run).max() <= 1e-4 * max(1.0, run.max())


def test_single_particle_driver_run():
    """DIEL=T alone (spoptics) with TMCOEF=T and SPEC=T: sp_optics.dat and the
    tm_coef files (k, band indices, then the same sorted rows), the
    single-particle spectra, and no BSE or polarised files (CPOL=F)."""
    say("case driver single particle")
    if not os.path.exists(DRIVER):
        pytest.fail("wtb_amd/wtb_bse is not built (make -C wtb_amd/fortran)")
    cards = ["DIEL= T", "TMCOEF= T", "SPEC= T", "SYSDIM= \"2D\"", "NGX= 12", "NGY= 12", "NGZ= 1", "NBANDSC= 1",
             "NBANDSV= 1", "CSHIFT= 0.05", "ENSPECI= 3.0", "ENSPECF= 7.0", "NESPEC= 400"]
    d, m = run_driver(cards)
    for absent in ("bse_opt_diel.dat", "log_bse-diel.dat", "bse_diel_xx.dat", "sp_diel-pol_x.dat"):
        assert not os.path.exists(os.path.join(d, absent)), absent
    sp = load(os.path.join(d, "sp_optics.dat"))
    assert sp.shape == (144, 7)
    tm = np.loadtxt(os.path.join(d, "tm_coef.dat"), skiprows=4, ndmin=2)
    assert tm.shape == (144, 13)
    assert np.abs(tm[:, 6:] - sp).max() <= 1.01e-4                    # F10.4 against F15.4
    assert np.all(tm[:, 3] == 1) and np.all(tm[:, 4] == 2) and np.all(tm[:, 5] == 1)   # nocp, c, v
    k = T.mp_mesh(12, 12, 1, m.rlat)
    kset = {tuple(np.round(x, 4)) for x in k.astype(np.float32).astype(float)}
    assert all(tuple(np.round(r, 4)) in kset for r in tm[:, :3])
    vc = abs(np.dot(np.cross(m.rlat[0], m.rlat[1]), m.rlat[2]))
    got = load(os.path.join(d, "sp_diel_xx.dat"))
    exp = O.spectrum(sp[:, 0], sp[:, 1], True, (12, 12, 1), vc, 3.0, 7.0, 400, 0.05, renorm=True)
    assert np.abs(got[:, 2] - np.float32(2.0 * exp[:, 2])).max() < 1e-6 * max(1.0, 2 * np.abs(exp[:, 2]).max())


def test_bz_activity_driver_run():
    """OPT_BZ=T alone (spoptpolbz): bz_act_{x,y,z,sp,sm}.dat and
    dichroism.dat per mesh point against the oracle sums."""
    say("case driver opt bz")
    if not os.path.exists(DRIVER):
        pytest.fail("wtb_amd/wtb_bse is not built (make -C wtb_amd/fortran)")
    cards = ["OPT_BZ= T", "NGX= 6", "NGY= 6", "NGZ= 1", "NBANDSC= 1", "NBANDSV= 1", "CSHIFT= 0.05"]
    d, m = run_driver(cards)
    k = T.mp_mesh(6, 6, 1, m.rlat)
    om = O.Model(m)
    r = O.hk_eig(om, k, 1, 1)
    ref = O.optics_bz(om, k, r["eig"], r["vec"], 1, 1, 0.05)
    for a, tag in enumerate(("x", "y", "z", "sp", "sm")):
        got = load(os.path.join(d, f"bz_act_{tag}.dat"))
        assert got.shape == (36, 4)
        assert np.abs(got[:, :3] - k.astype(np.float32)).max() < 1e-6
        assert np.abs(got[:, 3] - ref[:, a]).max() <= 1e-9 * max(1.0, np.abs(ref[:, a]).max()), tag
    dic = load(os.path.join(d, "dichroism.dat"))
    assert np.abs(dic[:, 3] - (ref[:, 3] - ref[:, 4]) / (ref[:, 3] + ref[:, 4])).max() < 1e-9


def test_log_file(honeycomb_run):
    """log.dat (CALC_DATA, here the run folder): the cards in param_out's
    layout (module_input.F90:422-556), the Begin line, the finished stages
    in main.F90's order, the End line."""
    d, m = honeycomb_run
    say("case driver log.dat")
    with open(os.path.join(d, "log.dat")) as fh:
        text = fh.read()
    lines = [ln.strip() for ln in text.splitlines()]
    for card in ("NTHREADS= 1", "SYSDIM= 2D", "BSE= T", "BSET= F", "SPEC= T", "PP_ONLY= F", "BSE_WF= T",
                 "DTDIAG= T", "CPOL= T", "NGX= 12", "NGY= 12", "NGZ= 1", "NBANDSC= 1", "NBANDSV= 1",
                 "ENSPECI=   3.0000", "ENSPECF=   7.0000", "NESPEC= 400", "COULOMB_POT= V2DRK",
                 "CSHIFT=   0.0500", "R_0=  10.0000", "EXC_WF_F= 4"):
        assert card in lines, card
    order = ["Begin", "BSE and Single particle dielectric calculation finished", "BSE dielectric properties calculated",
             "BSE absorption spectrum with light polarization calculated",
             "Single particle dielectric properties calculated",
             "Single particle absorption spectrum with light polarization calculated", "End"]
    pos = [text.index(x) for x in order]
    assert pos == sorted(pos)


def test_exciton_wavefunction_files(honeycomb_run):
    """BSE_WF=T: exc_wf_1..4.dat hold (nocp, k, c and v band, Re, Im) per
    state; the amplitudes' |psi|^2 summed over each degenerate level inside
    1..4 equal the diagonal of the oracle's projector on that level."""
    d, m = honeycomb_run
    say("case driver exciton wavefunctions")
    mesh, k, ref, w, X, dip = oracle_case(m)
    levels = np.split(np.arange(144), np.where(np.diff(w) > 1e-6)[0] + 1)
    inside = [lv for lv in levels if lv[-1] < 4]
    assert inside
    rows = {}
    for n in range(1, 5):
        f = os.path.join(d, f"exc_wf_{n}.dat")
        a = load(f)
        assert a.shape == (144, 8)
        assert np.all(a[:, 0] == 1) and np.all(a[:, 4] == 2) and np.all(a[:, 5] == 1)
        assert np.abs(a[:, 1:4] - k).max() <= 5.1e-5
        with open(f) as fh:
            head = fh.read().splitlines()[1].split()
        assert abs(float(head[-1]) - w[n - 1]) < 1e-8
        rows[n] = a[:, 6] + 1j * a[:, 7]
    for lv in inside:
        got = sum(np.abs(rows[i + 1]) ** 2 for i in lv)
        want = (np.abs(X[:, lv]) ** 2).sum(1)
        assert np.abs(got - want).max() < 1e-3 * want.max(), lv
