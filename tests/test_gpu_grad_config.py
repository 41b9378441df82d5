"""Gradients at the size of the differentiation workload: the seeded
22-orbital hexagonal model on a 48x48 mesh, nv = nc = 2 (N = 9216), FP64,
LOBPCG eigenpairs.

* lowest 8 exciton energies, random weights: r0, eps (EDIEL), the scissor and
  two Hermitian pairs of H(R) entries against central differences at step 1e-5;
* spectrum loss over a 1000-point grid against a seeded synthetic target
  (textbook convention): r0, eps and the scissor against a five-point stencil.
  The strengths depend on the eigenvectors, which LOBPCG delivers to its
  residual tolerance (1e-12 here), so the stencil uses step 1e-4.

Each value prints with its timing; the adjoint calls are timed too.
"""
import sys
import time

import numpy as np
import pytest

from conftest import hex_model
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu

MESH = (48, 48, 1)
NEV = 8
COUL = np.array([40.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0])
CSHIFT, GAMMA = 0.05, 0.05
C180 = float(np.float32(180.90))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def coul_dict(c):
    return dict(kind="V2DRK", r0=c[0], ediel=tuple(c[1:4]), ez=c[4], w=c[5], lc=c[6], ktol=0.001)


class Pipe:
    def __init__(self, conv):
        self.m = hex_model(MESH)
        self.k = T.mp_mesh(*MESH, self.m.rlat)
        self.ctx = hip.Context(0)
        self.ctx.set_kpoints(self.k, MESH)
        self.conv = conv
        self.cpre = C180 / (np.prod(MESH) * abs(np.linalg.det(self.m.rlat)))

    def run(self, coul=COUL, scissor=None, hop=None, tol=1e-12, vectors=False):
        m = T.TBModel(**{**self.m.__dict__})
        if hop is not None:
            m.hop = hop
        if scissor is not None:
            m.scissor = scissor
        self.ctx.set_tb_model(m)
        self.ctx.hk_eig(2, 2, host=False)
        self.ctx.bse_build(coul_dict(coul), hip.FP64, self.conv)
        E, X, rep = self.ctx.bse_solve(NEV, tol=tol, max_it=400, vectors=vectors)
        assert rep["nconv"] == NEV, rep
        return E, X

    def spectrum_loss(self, omega, target, **kw):
        E, X = self.run(vectors=True, **kw)
        d = self.ctx.optics(CSHIFT)[:, 0]
        f = np.abs(X.conj().T @ d) ** 2
        x = omega[:, None] - E[None, :]
        eps = self.cpre * (f[None, :] * GAMMA / (x * x + GAMMA * GAMMA)).sum(axis=1)
        return ((eps - target) ** 2).sum(), eps


def partner(m, r, a, b):
    rv = np.asarray(m.rvec)
    j = int(np.where((rv == -rv[r]).all(axis=1))[0][0])
    return j, b, a


def fd5(fun, h):
    return (-fun(2 * h) + 8 * fun(h) - 8 * fun(-h) + fun(-2 * h)) / (12 * h)


def test_energy_gradients_48x48():
    say("case energy gradients 48x48")
    p = Pipe(hip.CONV_REFERENCE)
    E, _ = p.run()
    say("N", p.ctx.N, "E", E)
    w = np.random.default_rng(5).uniform(0.5, 1.5, NEV)
    t = time.perf_counter()
    g = p.ctx.grad_energies(w)
    say("grad_energies (all H(R) entries) s", time.perf_counter() - t)
    H = 1e-5
    for name, i in (("r0", 0), ("ediel", 2)):
        h = H * abs(COUL[i])
        cp, cm = COUL.copy(), COUL.copy()
        cp[i] += h
        cm[i] -= h
        if name == "ediel":            # EDIEL enters as the mean of the three; move all
            cp[1:4] = COUL[1:4] + h
            cm[1:4] = COUL[1:4] - h
        fd = (w @ p.run(coul=cp)[0] - w @ p.run(coul=cm)[0]) / (2 * h)
        got = g["coulomb"][name] if name == "r0" else sum(g["coulomb"][k] for k in ("ediel_t", "ediel", "ediel_b"))
        say(name, "fd", fd, "adjoint", got)
        assert abs(got - fd) <= 1e-5 * abs(fd)
    fd = (w @ p.run(scissor=H)[0] - w @ p.run(scissor=-H)[0]) / (2 * H)
    say("scissor fd", fd, "adjoint", g["scissor"])
    assert abs(g["scissor"] - fd) <= 1e-5 * abs(fd)
    gh = g["hop"]
    scale = np.abs(gh).max()
    for r, a, b in ((0, 2, 15), (5, 10, 12)):
        j, bb, aa = partner(p.m, r, a, b)
        vals = []
        for s in (1, -1):
            hop = p.m.hop.copy()
            hop[r, a, b] += s * H
            hop[j, bb, aa] += s * H
            vals.append(w @ p.run(hop=hop)[0])
        fd = (vals[0] - vals[1]) / (2 * H)
        ex = gh[r, a, b].real + gh[j, bb, aa].real
        say("hop", r, a, b, "fd", fd, "adjoint", ex)
        assert abs(ex - fd) <= 1e-5 * (abs(fd) + 1e-3 * scale)
    p.ctx.close()


def test_spectrum_gradients_48x48():
    say("case spectrum gradients 48x48")
    p = Pipe(hip.CONV_TEXTBOOK)
    E, X = p.run(vectors=True)
    omega = np.linspace(E[0] - 0.3, E[-1] + 0.3, 1000)
    L0, eps = p.spectrum_loss(omega, np.zeros_like(omega))
    target = eps * np.random.default_rng(13).uniform(0.7, 1.2, omega.size)
    p.run(vectors=True)
    t = time.perf_counter()
    g = p.ctx.grad_spectrum(omega, target, "x", CSHIFT, GAMMA, hop=True)
    say("grad_spectrum (all H(R) entries) s", time.perf_counter() - t, "loss", g["loss"],
        "gauge residual", g["gauge_residual"])
    L, _ = p.spectrum_loss(omega, target)
    assert abs(g["loss"] - L) <= 1e-9 * L
    for name, i in (("r0", 0), ("ediel", 2), ("scissor", None)):
        if name == "scissor":
            fun = lambda d: p.spectrum_loss(omega, target, scissor=d)[0]
            fd = fd5(fun, 1e-4)
            got = g["scissor"]
        else:
            def fun(d, i=i, name=name):
                c = COUL.copy()
                if name == "ediel":
                    c[1:4] = COUL[1:4] + d
                else:
                    c[i] += d
                return p.spectrum_loss(omega, target, coul=c)[0]
            fd = fd5(fun, 1e-4 * abs(COUL[i]))
            got = g["coulomb"][name] if name == "r0" else sum(g["coulomb"][k]
                                                               for k in ("ediel_t", "ediel", "ediel_b"))
        say(name, "fd", fd, "adjoint", got)
        assert abs(got - fd) <= 1e-5 * abs(fd)
    p.ctx.close()
