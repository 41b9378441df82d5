"""Adjoint of the spectrum loss (wtb_grad_spectrum) against central finite
differences of the GPU FP64 pipeline (relative agreement 1e-5), and its
forward spectrum against the oracle's optdiel restatement.

L = sum_j (eps2(w_j) - t_j)^2 on a 40-point grid over the 8 lowest states of
the 22-orbital hexagonal model (8x8 mesh, nv = nc = 2, N = 256), anisotropic
V2DRK screening, Lorentzian width 0.05 eV, dipole broadening 0.05 eV.  The
finite differences rebuild everything (bands, dipoles, BSE matrix, solve) and
evaluate the loss from the oscillator strengths in numpy, so they share no
code with the adjoint beyond the forward kernels that the parity tests pin.
Hopping parameters are perturbed in Hermitian pairs as in the energy-gradient
tests.
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
COUL = np.array([30.0, 1.3, 2.0, 1.7, 1.5, 0.05, 4.0])
CSHIFT, GAMMA = 0.05, 0.05
H = 1e-5
C180 = float(np.float32(180.90))     # float32 literal of the reference (compat mode)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def coul_dict(c):
    return dict(kind="V2DRK", r0=c[0], ediel=tuple(c[1:4]), ez=c[4], w=c[5], lc=c[6], ktol=0.001)


def vcell(rlat):
    return abs(np.linalg.det(np.asarray(rlat, float)))


class Pipe:
    def __init__(self, conv):
        self.m = hex_model(MESH)
        self.k = T.mp_mesh(*MESH, self.m.rlat)
        self.ctx = hip.Context(0)
        self.ctx.set_kpoints(self.k, MESH)
        self.conv = conv
        self.cpre = C180 / (np.prod(MESH) * vcell(self.m.rlat))

    def run(self, coul=COUL, scissor=None, hop=None, nev=NEV):
        m = T.TBModel(**{**self.m.__dict__})
        if hop is not None:
            m.hop = hop
        if scissor is not None:
            m.scissor = scissor
        self.ctx.set_tb_model(m)
        self.ctx.hk_eig(2, 2)
        self.ctx.bse_build(coul_dict(coul), hip.FP64, self.conv)
        E, X, _ = self.ctx.bse_solve(nev)
        d = self.ctx.optics(CSHIFT)[:, 0]
        a = X.conj().T @ d if self.conv == hip.CONV_TEXTBOOK else X.T @ d
        return E, np.abs(a) ** 2

    def eps2(self, E, f, omega):
        x = omega[:, None] - E[None, :]
        return self.cpre * (f[None, :] * GAMMA / (x * x + GAMMA * GAMMA)).sum(axis=1)

    def loss(self, omega, target, **kw):
        E, f = self.run(**kw)
        return ((self.eps2(E, f, omega) - target) ** 2).sum()


def make_case(conv):
    p = Pipe(conv)
    E12, _ = p.run(nev=12)
    assert np.diff(E12[:NEV + 1]).min() > 1e-6, "test needs a non-degenerate lowest spectrum"
    E, f = p.run()
    omega = np.linspace(E[0] - 0.1, E[-1] + 0.1, 40)
    rng = np.random.default_rng(11)
    target = p.eps2(E, f, omega) * rng.uniform(0.6, 1.2, omega.size)
    return p, omega, target, E, f


def partner(m, r, a, b):
    rv = np.asarray(m.rvec)
    j = int(np.where((rv == -rv[r]).all(axis=1))[0][0])
    return j, b, a


def hop_expect(m, g, r, a, b, imag):
    j, bb, aa = partner(m, r, a, b)
    if (j, bb, aa) == (r, a, b):
        return g[r, a, b].real
    return g[r, a, b].imag - g[j, bb, aa].imag if imag else g[r, a, b].real + g[j, bb, aa].real


def fd5(fun, h):
    """five-point central difference, error O(h^4)"""
    return (-fun(2 * h) + 8 * fun(h) - 8 * fun(-h) + fun(-2 * h)) / (12 * h)


def check_scalars(p, omega, target, g, names):
    """Scalar parameters by the five-point stencil at step 1e-4 (relative for
    the screening parameters): the spectrum's third derivative in the
    scissor is large (the dipole factor dE / (dE^2 + eta^2) curves on the
    scale eta for the smallest transition energies), so the two-point
    difference at 1e-5 is off by ~4e-6, and the smallest screening gradients
    are ~1e-4 of the loss, where a 1e-5 step leaves rounding noise of ~1e-5."""
    for name in names:
        if name == "scissor":
            fd = fd5(lambda d: p.loss(omega, target, scissor=d), 1e-4)
            got = g["scissor"]
        else:
            i = hip.COULOMB_PARAMS.index(name)

            def at(d):
                cc = COUL.copy()
                cc[i] += d
                return p.loss(omega, target, coul=cc)

            fd = fd5(at, 1e-4 * max(1.0, abs(COUL[i])))
            got = g["coulomb"][name]
        say("param", name, "fd", fd, "adjoint", got)
        assert abs(got - fd) <= 1e-5 * abs(fd), (name, got, fd)


@pytest.fixture(scope="module")
def textbook():
    p, omega, target, E, f = make_case(hip.CONV_TEXTBOOK)
    p.run()
    g = p.ctx.grad_spectrum(omega, target, "x", CSHIFT, GAMMA, hop=True)
    say("textbook loss", g["loss"], "gauge residual", g["gauge_residual"])
    return p, omega, target, E, f, g


def test_spectrum_forward_matches_oracle(textbook):
    p, omega, target, E, f, g = textbook
    say("case spectrum forward")
    eps = p.eps2(E, f, omega)
    assert np.abs(g["eps2"] - eps).max() <= 1e-10 * np.abs(eps).max()
    assert abs(g["loss"] - ((eps - target) ** 2).sum()) <= 1e-10 * g["loss"]
    # the oracle's optdiel: width = avgsme(E) + sme, grid e0 + (e1 - e0) j / (npts - 1)
    avg = 10.0 * (np.diff(E).sum() / E.size)
    ref = O.spectrum(E, f, False, MESH, vcell(p.m.rlat), omega[0], omega[-1], omega.size, GAMMA - avg,
                     renorm=False)
    assert np.abs(ref[:, 0] - omega).max() < 1e-12
    assert np.abs(ref[:, 2] - g["eps2"]).max() <= 1e-9 * np.abs(ref[:, 2]).max()


def test_spectrum_grad_textbook_scalars_fd(textbook):
    p, omega, target, E, f, g = textbook
    say("case spectrum grad textbook scalars")
    assert g["gauge_residual"] < 1e-8 * max(1.0, np.abs(g["hop"]).max())
    check_scalars(p, omega, target, g, ["r0", "ediel_t", "w", "scissor"])


ENTRIES = [(0, 3, 3, False), (0, 2, 15, False), (0, 2, 15, True), (5, 10, 12, False), (5, 10, 12, True),
           (20, 0, 21, True)]


def test_spectrum_grad_textbook_hoppings_fd(textbook):
    p, omega, target, E, f, g = textbook
    say("case spectrum grad textbook hoppings")
    gh = g["hop"]
    scale = np.abs(gh).max()
    for r, a, b, imag in ENTRIES:
        j, bb, aa = partner(p.m, r, a, b)
        vals = []
        for s in (1, -1):
            hop = p.m.hop.copy()
            d = (1j if imag else 1.0) * s * H
            hop[r, a, b] += d
            if (j, bb, aa) != (r, a, b):
                hop[j, bb, aa] += np.conj(d)
            vals.append(p.loss(omega, target, hop=hop))
        fd = (vals[0] - vals[1]) / (2 * H)
        ex = hop_expect(p.m, gh, r, a, b, imag)
        say("entry", r, a, b, imag, "fd", fd, "adjoint", ex)
        assert abs(ex - fd) <= 1e-5 * (abs(fd) + 1e-3 * scale)


def test_spectrum_grad_reference_convention_fd():
    say("case spectrum grad reference convention")
    p, omega, target, E, f = make_case(hip.CONV_REFERENCE)
    # the strengths are those of wtb_oscillator
    act = p.ctx.oscillator(NEV, "x", "x")
    assert np.abs(act[:NEV] - f).max() <= 1e-12 * max(1.0, f.max())
    g = p.ctx.grad_spectrum(omega, target, "x", CSHIFT, GAMMA, hop=False)
    assert abs(g["loss"] - ((p.eps2(E, f, omega) - target) ** 2).sum()) <= 1e-10 * g["loss"]
    check_scalars(p, omega, target, g, ["r0", "ediel", "scissor"])
    with pytest.raises(hip.WTBError) as e:
        p.ctx.grad_spectrum(omega, target, "x", CSHIFT, GAMMA, hop=True)
    assert e.value.code == -7          # WTB_ERR_UNSUPPORTED
    p.ctx.close()


def test_spectrum_torch_function_and_fit():
    torch = pytest.importorskip("torch")
    from wtb_amd import fit as F
    say("case spectrum torch harness")
    m = hex_model(MESH)
    prob = F.ExcitonProblem(m, MESH, 2, 2, nev=NEV, convention=hip.CONV_TEXTBOOK)
    E = prob.solve(COUL, 0.0, m.hop)
    omega = np.linspace(E[0] - 0.1, E[-1] + 0.1, 40)
    spec = dict(omega=omega, target=np.zeros_like(omega), comp="x", cshift=CSHIFT, gamma=GAMMA)
    spec["target"] = prob.spectrum_grad(spec, hop=False)["eps2"].copy()
    start = COUL.copy()
    start[0] = 27.0
    coul = torch.tensor(start, requires_grad=True)
    sc = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    hre = torch.tensor(m.hop.real.copy(), requires_grad=True)
    him = torch.tensor(m.hop.imag.copy(), requires_grad=True)
    L = F.SpectrumLoss.apply(prob, spec, coul, sc, hre, him)
    (3.0 * L).backward()
    g = prob.spectrum_grad(spec, hop=True)     # same solve state
    assert abs(float(L) - g["loss"]) <= 1e-12 * g["loss"]
    assert np.allclose(coul.grad.numpy(), [3.0 * g["coulomb"][k] for k in hip.COULOMB_PARAMS], rtol=1e-12, atol=0)
    assert abs(float(sc.grad) - 3.0 * g["scissor"]) <= 1e-12 * abs(g["scissor"]) + 1e-300
    assert np.allclose(hre.grad.numpy(), 3.0 * g["hop"].real, rtol=1e-12, atol=1e-300)
    out = F.fit_spectrum(prob, spec, torch.tensor(start), torch.tensor(0.0), hre.detach(), him.detach(),
                         fit_coul=(0,), steps=25)
    say("spectrum fit history", out["history"][:3], out["history"][-3:], "r0", float(out["coul"][0]))
    assert abs(float(out["coul"][0]) - COUL[0]) < 1e-3 * COUL[0]
    prob.close()


def oracle_loss(m, k, coul, omega, target, cpre, scissor=0.0):
    """The reference-convention spectrum loss entirely from the oracle:
    bands, BSE matrix, dense eigenpairs, dipoles, strengths |x^T d|^2."""
    mm = T.TBModel(**{**m.__dict__})
    mm.scissor = scissor
    om = O.Model(mm)
    ref = O.hk_eig(om, k, 2, 2)
    Hm = O.bse_matrix(k, ref["eig"], ref["vec"], O.coulomb(**coul_dict(coul)), MESH, mm.rlat, 2, 2)
    E, X = np.linalg.eigh(Hm)
    E, X = E[:NEV], X[:, :NEV]
    d = O.dipoles(om, k, ref["eig"], ref["vec"], 2, 2, CSHIFT)[0]
    f = np.abs(X.T @ d) ** 2
    x = omega[:, None] - E[None, :]
    eps = cpre * (f[None, :] * GAMMA / (x * x + GAMMA * GAMMA)).sum(axis=1)
    return ((eps - target) ** 2).sum()


def test_spectrum_grad_against_oracle_fd():
    """Reference convention: the GPU adjoint against finite differences of
    the oracle's own pipeline (r0 and the scissor, five-point stencil)."""
    say("case spectrum grad vs oracle fd")
    p, omega, target, E, f = make_case(hip.CONV_REFERENCE)
    g = p.ctx.grad_spectrum(omega, target, "x", CSHIFT, GAMMA, hop=False)
    L0 = oracle_loss(p.m, p.k, COUL, omega, target, p.cpre)
    assert abs(L0 - g["loss"]) <= 1e-9 * L0

    def at_r0(dv):
        c = COUL.copy()
        c[0] += dv
        return oracle_loss(p.m, p.k, c, omega, target, p.cpre)

    fd = fd5(at_r0, 1e-4 * COUL[0])
    say("r0 oracle fd", fd, "adjoint", g["coulomb"]["r0"])
    assert abs(g["coulomb"]["r0"] - fd) <= 1e-5 * abs(fd)
    fd = fd5(lambda dv: oracle_loss(p.m, p.k, COUL, omega, target, p.cpre, scissor=dv), 1e-4)
    say("scissor oracle fd", fd, "adjoint", g["scissor"])
    assert abs(g["scissor"] - fd) <= 1e-5 * abs(fd)
    p.ctx.close()
