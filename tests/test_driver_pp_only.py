"""PP_ONLY=T (main.F90:79): the Fortran driver skips every calculation and
recomputes the spectra, optical constants and lifetimes from the files
already in OUTPUT.  Runs without a GPU: the input files are written here
(synthetic exciton energies and strengths with the reference's 7F15.4
rows), and the outputs are checked against the oracle spectrum
(optdiel/avgsme/imagrenorm) of those files and against log.dat."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from wtb_amd import tbmodel as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "wtb_amd", "wtb_bse")


def write_rows(path, header, rows):
    with open(path, "w") as f:
        f.write(" " + header + "\n")
        for r in rows:
            f.write("".join(f"{x:15.4f}" for x in r) + "\n")


def test_pp_only_recomputes_spectra_from_the_written_files():
    if not os.path.exists(DRIVER):
        pytest.skip("wtb_amd/wtb_bse is not built")
    d = tempfile.mkdtemp(prefix="wtbpp")
    m = T.honeycomb_two_band()
    T.write_tb_file(m, os.path.join(d, "tb.dat"))
    rng = np.random.default_rng(5)
    n = 144
    e = np.sort(rng.uniform(3.2, 6.5, n))
    act = np.abs(rng.standard_normal((n, 6))) * np.array([1, 1, 0.01, 0.1, 0.0, 0.0])
    act[100:] = 0.0
    e[100:] = 0.0                       # rows beyond nev hold zeros, as the driver writes them
    write_rows(os.path.join(d, "bse_opt_diel.dat"), "# exciton energy xx yy zz xy xz yz", np.c_[e, act])
    sp_e = np.sort(rng.uniform(3.5, 7.0, n))
    sp = np.abs(rng.standard_normal((n, 6)))
    write_rows(os.path.join(d, "sp_optics.dat"), "# energy xx yy zz xy xz yz", np.c_[sp_e, sp])
    write_rows(os.path.join(d, "sp_optics-pol.dat"), "# energy xx yy zz sp sm", np.c_[sp_e, sp[:, :5]])
    cards = [f'PARAMS_FILE= "{d}/tb.dat"', f'OUTPUT= "{d}/"', f'CALC_DATA= "{d}/"', "PP_ONLY= T", "BSE= T",
             "SPEC= T", 'SYSDIM= "2D"', "NGX= 12", "NGY= 12", "NGZ= 1", "CSHIFT= 0.05", "ENSPECI= 3.0",
             "ENSPECF= 7.0", "NESPEC= 300"]
    r = subprocess.run([DRIVER], input="\n".join(cards) + "\n", capture_output=True, text=True, timeout=120, cwd=d)
    assert r.returncode == 0, r.stdout + r.stderr
    # no calculation ran: none of the BSE path's own outputs
    for absent in ("log_bse-diel.dat", "bse_opt_diel-pol.dat", "exc_wf_1.dat"):
        assert not os.path.exists(os.path.join(d, absent)), absent
    out = np.loadtxt(os.path.join(d, "bse_opt_diel.dat"), comments="#", ndmin=2)
    vc = abs(np.dot(np.cross(m.rlat[0], m.rlat[1]), m.rlat[2]))
    for name, col, diag in (("xx", 1, True), ("yy", 2, True), ("xy", 4, False)):
        got = np.loadtxt(os.path.join(d, f"bse_diel_{name}.dat"), ndmin=2)
        exp = O.spectrum(out[:, 0], out[:, col], diag, (12, 12, 1), vc, 3.0, 7.0, 300, 0.05, renorm=True)
        assert got.shape == (300, 3)
        scale = max(1.0, 2.0 * np.abs(exp[:, 1:]).max())
        assert np.abs(got[:, 1] - np.float32(2.0 * exp[:, 1])).max() < 1e-6 * scale, name
        assert np.abs(got[:, 2] - np.float32(2.0 * exp[:, 2])).max() < 1e-6 * scale, name
    spx = np.loadtxt(os.path.join(d, "sp_diel_xx.dat"), ndmin=2)
    spf = np.loadtxt(os.path.join(d, "sp_optics.dat"), comments="#", ndmin=2)
    exp = O.spectrum(spf[:, 0], spf[:, 1], True, (12, 12, 1), vc, 3.0, 7.0, 300, 0.05, renorm=True)
    assert np.abs(spx[:, 2] - np.float32(2.0 * exp[:, 2])).max() < 1e-6 * max(1.0, 2 * np.abs(exp[:, 2]).max())
    for f in ("bse_absorption_coef.dat", "exciton_lifetime.dat", "sp_absorption_coef.dat"):
        assert os.path.exists(os.path.join(d, f)), f
    with open(os.path.join(d, "log.dat")) as fh:
        log = fh.read()
    assert "PP_ONLY= T" in log and "Begin" in log and "End" in log
    assert "BSE dielectric properties calculated" in log
    assert "calculation finished" not in log
