"""Adjoint of the exciton energies (wtb_grad_energies) against central finite
differences (step 1e-5) of the GPU FP64 pipeline and of the oracle pipeline,
plus the torch.autograd.Function harness.

L = sum_n w_n E_n over the 8 lowest states with random weights on the 22-orbital
hexagonal model (8x8 mesh, nv = nc = 2, N = 256), anisotropic V2DRK screening.
Hopping parameters are perturbed in Hermitian pairs (H(R)_ab with H(-R)_ba), the
directions in which H(k) stays Hermitian: the FD derivative along such a pair
is Re g(R)_ab + Re g(-R)_ba (real part) or Im g(R)_ab - Im g(-R)_ba (imaginary).
"""
import sys

import numpy as np
import pytest

from conftest import hex_model
from oracle import oracle as O
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu

MESH = (8, 8, 1)
NEV = 8
COUL = np.array([30.0, 1.3, 2.0, 1.7, 1.5, 0.05, 4.0])    # r0, EDIEL_T, EDIEL, EDIEL_B, EDIEL_Z, W_COUL, LC
H = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def coul_dict(c, kind="V2DRK"):
    return dict(kind=kind, r0=c[0], ediel=tuple(c[1:4]), ez=c[4], w=c[5], lc=c[6], ktol=0.001)


class Pipe:
    def __init__(self, conv=hip.CONV_REFERENCE, kind="V2DRK"):
        self.m = hex_model(MESH)
        self.k = T.mp_mesh(*MESH, self.m.rlat)
        self.ctx = hip.Context(0)
        self.ctx.set_kpoints(self.k, MESH)
        self.conv, self.kind = conv, kind

    def energies(self, coul=COUL, scissor=None, hop=None):
        m = T.TBModel(**{**self.m.__dict__})
        if hop is not None:
            m.hop = hop
        if scissor is not None:
            m.scissor = scissor
        self.ctx.set_tb_model(m)
        self.ctx.hk_eig(2, 2)
        self.ctx.bse_build(coul_dict(coul, self.kind), hip.FP64, self.conv)
        E, _, _ = self.ctx.bse_solve(NEV)
        return E


def partner(m, r, a, b):
    """index of -R and the transposed entry"""
    rv = np.asarray(m.rvec)
    j = int(np.where((rv == -rv[r]).all(axis=1))[0][0])
    return j, b, a


@pytest.fixture(scope="module")
def base():
    p = Pipe()
    E = p.energies()
    gaps = np.diff(E)
    assert gaps.min() > 1e-6, "test needs a non-degenerate lowest spectrum"
    w = np.random.default_rng(7).uniform(0.5, 1.5, NEV)
    g = p.ctx.grad_energies(w)
    say("E", E, "grad", g["coulomb"], g["scissor"], g["efermi"], g["gauge_residual"])
    return p, w, g


def test_grad_is_reproducible():
    """The per-workgroup partials are summed in a fixed order (no atomics):
    two calls on the same solve give the same gradient bit for bit."""
    say("case grad reproducible")
    p = Pipe()
    p.energies()
    w = np.random.default_rng(7).uniform(0.5, 1.5, NEV)
    g1 = p.ctx.grad_energies(w)
    g2 = p.ctx.grad_energies(w)
    assert list(g1["coulomb"].values()) == list(g2["coulomb"].values())
    assert g1["scissor"] == g2["scissor"] and g1["efermi"] == g2["efermi"]
    assert np.array_equal(g1["hop"], g2["hop"])
    p.ctx.close()


def test_grad_coulomb_parameters_fd(base):
    p, w, g = base
    say("case grad coulomb")
    names = hip.COULOMB_PARAMS
    fd, hs = [], []
    for i, name in enumerate(names):
        h = H * max(1.0, abs(COUL[i]))
        cp, cm = COUL.copy(), COUL.copy()
        cp[i] += h
        cm[i] -= h
        fd.append((w @ p.energies(cp) - w @ p.energies(cm)) / (2 * h))
        hs.append(h)
    fd, hs = np.array(fd), np.array(hs)
    got = np.array([g["coulomb"][n] for n in names])
    say("fd", fd, "adjoint", got, "|diff|", np.abs(got - fd))
    # 1e-5 relative, plus the central difference's own rounding floor: the
    # FP64 solve leaves ~3e-14 eV of noise on w.E, i.e. up to ~4e-10 on a
    # derivative at step 4e-5 (seen on every component, whatever its size);
    # 1e-13 / (2 h) bounds it.
    assert np.all(np.abs(got - fd) <= 1e-5 * np.abs(fd) + 1e-13 / (2 * hs)), (got, fd)


def test_grad_scissor_efermi_gauge(base):
    p, w, g = base
    say("case grad scissor")
    assert abs(g["scissor"] - w.sum()) < 1e-10 * w.sum()
    fd = (w @ p.energies(scissor=H) - w @ p.energies(scissor=-H)) / (2 * H)
    assert abs(fd - w.sum()) < 1e-6 * w.sum()
    assert abs(g["efermi"]) < 1e-9 * w.sum()
    assert g["gauge_residual"] < 1e-9


def hop_fd(p, w, r, a, b, imag):
    j, bb, aa = partner(p.m, r, a, b)
    out = []
    for s in (1, -1):
        hop = p.m.hop.copy()
        d = (1j if imag else 1.0) * s * H
        hop[r, a, b] += d
        if (j, bb, aa) != (r, a, b):
            hop[j, bb, aa] += np.conj(d)
        out.append(w @ p.energies(hop=hop))
    return (out[0] - out[1]) / (2 * H)


def hop_expect(m, g, r, a, b, imag):
    j, bb, aa = partner(m, r, a, b)
    if (j, bb, aa) == (r, a, b):
        return g[r, a, b].real
    return g[r, a, b].imag - g[j, bb, aa].imag if imag else g[r, a, b].real + g[j, bb, aa].real


ENTRIES = [(0, 3, 3, False), (0, 2, 15, False), (0, 2, 15, True), (5, 10, 12, False), (5, 10, 12, True),
           (20, 0, 21, True)]


def test_grad_hoppings_fd(base):
    p, w, g = base
    say("case grad hoppings")
    gh = g["hop"]
    assert gh.shape == (p.m.nvec, p.m.norb, p.m.norb)
    scale = np.abs(gh).max()
    for r, a, b, imag in ENTRIES:
        fd = hop_fd(p, w, r, a, b, imag)
        ex = hop_expect(p.m, gh, r, a, b, imag)
        say("entry", r, a, b, imag, "fd", fd, "adjoint", ex)
        assert abs(ex - fd) <= 1e-5 * (abs(fd) + 1e-3 * scale)


def test_grad_textbook_convention_fd():
    p = Pipe(conv=hip.CONV_TEXTBOOK)
    say("case grad textbook")
    E = p.energies()
    w = np.random.default_rng(8).uniform(0.5, 1.5, NEV)
    g = p.ctx.grad_energies(w)
    gh = g["hop"]
    scale = np.abs(gh).max()
    for r, a, b, imag in ENTRIES[1:4]:
        fd = hop_fd(p, w, r, a, b, imag)
        ex = hop_expect(p.m, gh, r, a, b, imag)
        assert abs(ex - fd) <= 1e-5 * (abs(fd) + 1e-3 * scale)
    h = H * COUL[0]
    cp, cm = COUL.copy(), COUL.copy()
    cp[0] += h
    cm[0] -= h
    fd = (w @ p.energies(cp) - w @ p.energies(cm)) / (2 * h)
    assert abs(g["coulomb"]["r0"] - fd) <= 1e-5 * abs(fd)


def oracle_L(m, k, w, coul, hop=None):
    mm = T.TBModel(**{**m.__dict__})
    if hop is not None:
        mm.hop = hop
    ref = O.hk_eig(O.Model(mm), k, 2, 2)
    Hm = O.bse_matrix(k, ref["eig"], ref["vec"], O.coulomb(**coul_dict(coul)), MESH, mm.rlat, 2, 2)
    return w @ np.linalg.eigvalsh(Hm)[:NEV]


def test_grad_against_oracle_fd(base):
    p, w, g = base
    say("case grad vs oracle fd")
    h = H * COUL[0]
    cp, cm = COUL.copy(), COUL.copy()
    cp[0] += h
    cm[0] -= h
    fd = (oracle_L(p.m, p.k, w, cp) - oracle_L(p.m, p.k, w, cm)) / (2 * h)
    assert abs(g["coulomb"]["r0"] - fd) <= 1e-5 * abs(fd)
    r, a, b = 5, 10, 12
    j, bb, aa = partner(p.m, r, a, b)
    vals = []
    for s in (1, -1):
        hop = p.m.hop.copy()
        hop[r, a, b] += s * H
        hop[j, bb, aa] += s * H
        vals.append(oracle_L(p.m, p.k, w, COUL, hop))
    fd = (vals[0] - vals[1]) / (2 * H)
    ex = hop_expect(p.m, g["hop"], r, a, b, False)
    assert abs(ex - fd) <= 1e-5 * (abs(fd) + 1e-3 * np.abs(g["hop"]).max())


def test_torch_function_and_fit():
    torch = pytest.importorskip("torch")
    from wtb_amd import fit as F
    say("case torch harness")
    m = hex_model(MESH)
    prob = F.ExcitonProblem(m, MESH, 2, 2, nev=NEV)
    coul = torch.tensor(COUL, dtype=torch.float64, requires_grad=True)
    sc = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    hre = torch.tensor(m.hop.real.copy(), requires_grad=True)
    him = torch.tensor(m.hop.imag.copy(), requires_grad=True)
    E = F.ExcitonEnergies.apply(prob, coul, sc, hre, him)
    w = torch.linspace(0.5, 1.5, NEV, dtype=torch.float64)
    (w * E).sum().backward()
    g = prob.grad(w.numpy())        # same solve state
    assert np.allclose(coul.grad.numpy(), [g["coulomb"][k] for k in hip.COULOMB_PARAMS], rtol=1e-12, atol=0)
    assert abs(float(sc.grad) - float(w.sum())) < 1e-9
    assert np.allclose(hre.grad.numpy(), g["hop"].real, rtol=1e-12, atol=1e-15)
    # fit r0 back from a perturbed start
    target = prob.solve(COUL, 0.0, m.hop)
    start = COUL.copy()
    start[0] = 22.0
    out = F.fit(prob, target, torch.tensor(start), torch.tensor(0.0), hre.detach(), him.detach(),
                fit_coul=(0,), steps=25)
    say("fit history", out["history"][:3], out["history"][-3:], "r0", float(out["coul"][0]))
    assert abs(float(out["coul"][0]) - COUL[0]) < 1e-3 * COUL[0]
    prob.close()
