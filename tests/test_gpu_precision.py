"""The mixed-precision pattern at full size on the 3-D bulk workload: the seeded
16-orbital cubic model on a 16x16x16 mesh, nv = nc = 3 (N = 36864), V3DL
screening with EDIEL = 6, lowest 128 eigenpairs.  The mixed path (FP32 store,
split-FP16 products, int8 refinement, FP64 Rayleigh-Ritz) must match the FP64
path to 0.1 meV on every energy; the absorption spectrum (2000 points over
0.5-6 eV, 0.05 eV Lorentzian, trace of the Cartesian strengths) is compared
in relative L2 norm.  This workload runs the wide block (144 columns: two
split-FP16 launches and three int8 launches per product) and a 432-wide
Rayleigh-Ritz problem.
"""
import sys
import time

import numpy as np
import pytest

from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_precision_pattern_cubic16():
    model, spec = T.workload("cubic16-16")
    mesh, nev = spec["mesh"], spec["nev"]
    say("case cubic16-16 precision pattern, nev", nev)
    k = T.mp_mesh(*mesh, model.rlat)
    ctx = hip.Context(0)
    ctx.set_tb_model(model)
    ctx.set_kpoints(k, mesh)
    ctx.hk_eig(spec["nv"], spec["nc"], host=False)
    assert ctx.N == 36864
    out = {}
    for prec, name in ((hip.MIXED, "mixed"), (hip.FP64, "fp64")):
        t = time.perf_counter()
        ctx.bse_build(spec["coulomb"], prec)
        E, _, rep = ctx.bse_solve(nev, vectors=False)
        dt = time.perf_counter() - t
        ctx.optics(spec["cshift"])
        f = sum(ctx.oscillator(nev, a)[:nev] for a in ("x", "y", "z"))
        out[name] = (E, f)
        say(name, "build + solve s", round(dt, 4), "iterations", rep["iterations"], "columns",
            rep["matvec_columns"], "max rel residual", rep["max_rel_residual"])
        assert rep["nconv"] == nev
    dE = np.abs(out["mixed"][0] - out["fp64"][0])
    say("max |dE| meV", dE.max() * 1e3)
    assert dE.max() < 1e-4
    e0, e1, npts = spec["spectrum"]
    omega = np.linspace(e0, e1, npts)

    def spectrum(E, f):
        x = omega[:, None] - E[None, :]
        return (f[None, :] * spec["cshift"] / (x * x + spec["cshift"] ** 2)).sum(axis=1)

    s64, smix = spectrum(*out["fp64"]), spectrum(*out["mixed"])
    rel = np.linalg.norm(smix - s64) / np.linalg.norm(s64)
    say("spectrum relative L2", rel)
    assert rel < 1e-4
    ctx.close()
