"""Per-k optical activity (OPT_BZ=T, spoptpolbz) on the device against the oracle."""
import sys

import numpy as np
import pytest

from conftest import hex_model
from oracle import oracle as O
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu


def test_optics_bz_matches_oracle():
    print("case optics bz", flush=True)
    sys.stdout.flush()
    mesh = (6, 6, 1)
    m = hex_model(mesh)
    k = T.mp_mesh(*mesh, m.rlat)
    with hip.Context(0) as ctx:
        ctx.set_tb_model(m)
        ctx.set_kpoints(k, mesh)
        ctx.hk_eig(2, 2, host=False)
        act = ctx.optics_bz(0.05)
    om = O.Model(m)
    r = O.hk_eig(om, k, 2, 2)
    ref = O.optics_bz(om, k, r["eig"], r["vec"], 2, 2, 0.05)
    assert act.shape == (36, 5)
    assert np.abs(act - ref).max() < 1e-10 * np.abs(ref).max()
