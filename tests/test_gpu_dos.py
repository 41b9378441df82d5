"""Density of states (DOS=T, dostool) on the device against the oracle."""
import sys

import numpy as np
import pytest

from conftest import hex_model
from oracle import oracle as O
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_dos_matches_oracle(mode):
    say("case dos", mode)
    mesh = (8, 8, 1)
    m = hex_model(mesh)
    k = T.mp_mesh(*mesh, m.rlat)
    ow = np.random.default_rng(7).uniform(0.0, 1.0, m.norb) if mode else None
    with hip.Context(0) as ctx:
        ctx.set_tb_model(m)
        ctx.set_kpoints(k, mesh)
        ctx.hk_eig(2, 2, host=False)
        en, d = ctx.dos(501, 0.08, orbweight=ow, spin_mode=mode)
    r = O.hk_eig(O.Model(m), k, 2, 2, full=True)
    en_r, d_r = O.dos(r["eall"], r["zall"], 0.08, 501, orbweight=ow, spin_mode=mode)
    assert np.abs(en - en_r).max() < 1e-10
    # energies agree to 1e-10 eV; a Gaussian of width 0.08 turns that into ~1e-8 relative
    assert np.abs(d - d_r).max() < 1e-8 * np.abs(d_r).max()
    if mode:
        assert np.abs(d[:, 1] + d[:, 2] - d[:, 0]).max() < 1e-12 * d[:, 0].max()


def test_dos_arguments():
    say("case dos arguments")
    mesh = (4, 4, 1)
    m = hex_model(mesh)
    with hip.Context(0) as ctx:
        ctx.set_tb_model(m)
        ctx.set_kpoints(T.mp_mesh(*mesh, m.rlat), mesh)
        with pytest.raises(hip.WTBError) as e:
            ctx.dos(101, 0.1)
        assert e.value.code == -6               # no bands yet
        ctx.hk_eig(1, 1, host=False)
        for bad in (dict(ndos=1, sigma=0.1), dict(ndos=11, sigma=0.0), dict(ndos=11, sigma=0.1, spin_mode=3)):
            with pytest.raises(hip.WTBError) as e:
                ctx.dos(**bad)
            assert e.value.code == -1
