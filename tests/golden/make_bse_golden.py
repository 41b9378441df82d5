"""Golden exciton energies and oscillator strengths of the BSE workloads.

Test-data generator, run in the development container (never on the GPU box,
never by the product).  For a named workload it builds the oracle's FP64 BSE
matrix (the reference's matrizelbse and assembly loop, module_bse.F90:24-83,
2392-2415, restated in oracle/wtb_oracle.c), takes its lowest ``nev``
eigenpairs from LAPACK (scipy.linalg.eigh, zheevr with subset_by_index, so no
part of this build's solver is involved) and the in-plane oscillator
strengths |sum_i X_i d_i|^2 of dielbsev (module_bse.F90:201-245) from the
oracle dipoles (optsp/hlm, module_optics.F90:12-167, broadening 0.05 eV).

The model is regenerated from its seed by wtb_amd.tbmodel; its fingerprint
(sha256 of the hoppings, lattice and Fermi energy) is stored so a test can
check that it solves the same matrix.

    python tests/golden/make_bse_golden.py honeycomb-24 hex22-60 hex22-90
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# workload -> number of lowest pairs kept
NEV = {"honeycomb-24": 576, "hex22-60": 64, "hex22-90": 64, "mos2-44": 100}

# The reference's own example 3 (examples/example-3/example_3.sh): MoS2 with
# spin-orbit coupling (tb_mos2.dat, norb = 22, 147 R vectors, six of
# Wigner-Seitz degeneracy 2), RK = 120 on the 2-D lattice -> 44 x 44 mesh,
# nv = nc = 2 (N = 7744), V2DT2 with LC = 8, CSHIFT = 0.08; 100 pairs, the
# reference's hard-coded SLEPc nev (module_bse.F90:2432).
MOS2_FILE = os.path.join(HERE, "tb_mos2.dat")
MOS2_SPEC = dict(mesh=(44, 44, 1), nv=2, nc=2,
                 coulomb=dict(kind="V2DT2", r0=1.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=8.0, ktol=0.001),
                 cshift=0.08)


def workload(name):
    from wtb_amd import tbmodel as T
    if name == "mos2-44":
        return T.read_tb_file(MOS2_FILE), dict(MOS2_SPEC)
    return T.workload(name)


def model_fingerprint(model):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(model.hop, np.complex128).tobytes())
    h.update(np.ascontiguousarray(model.rvec, np.int64).tobytes())
    h.update(np.ascontiguousarray(model.deg, np.int64).tobytes())
    h.update(np.ascontiguousarray(model.rlat, np.float64).tobytes())
    h.update(np.float64(model.efermi).tobytes())
    h.update(np.float64(model.scissor).tobytes())
    return h.hexdigest()


def golden_path(name):
    return os.path.join(HERE, f"bse_{name}.json")


def load(name):
    with open(golden_path(name)) as f:
        return json.load(f)


def make(name):
    import scipy
    import scipy.linalg as sl
    from oracle import oracle as O
    from wtb_amd import tbmodel as T

    model, spec = workload(name)
    mesh = spec["mesh"]
    nv, nc = spec["nv"], spec["nc"]
    nev = NEV[name]
    kpt = T.mp_mesh(*mesh, model.rlat)
    om = O.Model(model)
    t0 = time.perf_counter()
    r = O.hk_eig(om, kpt, nv, nc)
    assert r["status"] == 0
    N = kpt.shape[0] * nc * nv
    c = O.coulomb(**spec["coulomb"])
    h = O.bse_rows(kpt, r["eig"], r["vec"], c, mesh, model.rlat, nc, nv, 0, N)
    t1 = time.perf_counter()
    print(f"{name}: N={N} oracle matrix in {t1 - t0:.1f} s", flush=True)
    # h is row-major H; its transpose view is Fortran-ordered H^T = conj(H):
    # same eigenvalues, eigenvectors conj(X)
    w, z = sl.eigh(h.T, overwrite_a=True, check_finite=False, subset_by_index=[0, min(N, nev + 1) - 1],
                   driver="evr")
    del h
    t2 = time.perf_counter()
    print(f"{name}: zheevr lowest {len(w)} in {t2 - t1:.1f} s", flush=True)
    x = np.conj(z[:, :nev])
    d = O.dipoles(om, kpt, r["eig"], r["vec"], nc, nv, spec["cshift"])
    fx = O.oscillator(x, d[0], d[0])[:nev]
    fy = O.oscillator(x, d[1], d[1])[:nev]
    gaps = np.diff(w)
    out = {
        "workload": name, "N": int(N), "nev": int(nev), "mesh": list(mesh), "nv": nv, "nc": nc,
        "coulomb": spec["coulomb"], "cshift": spec["cshift"],
        "model_sha256": model_fingerprint(model),
        "flags": "COMPAT_F32|GAUGE_LAPACK",
        "eigenvalues_eV": [float(v) for v in w[:nev]],
        "next_eigenvalue_eV": float(w[nev]) if len(w) > nev else None,
        "min_spacing_eV": float(gaps.min()) if len(gaps) else None,
        "osc_x": [float(v) for v in fx], "osc_y": [float(v) for v in fy],
        "generator": "tests/golden/make_bse_golden.py: oracle FP64 matrix, scipy.linalg.eigh driver=evr",
        "lapack": f"scipy {scipy.__version__}, " + np.__version__,
    }
    with open(golden_path(name), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{name}: written {golden_path(name)}", flush=True)


if __name__ == "__main__":
    for nm in sys.argv[1:]:
        make(nm)
