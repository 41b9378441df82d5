"""PyTorch fitting harness over the HIP exciton path and its adjoint.

`ExcitonEnergies` is a torch.autograd.Function whose forward runs the GPU
pipeline (H(k) eigensolves, BSE assembly, lowest eigenpairs) for given
screening parameters, scissor shift and TB hoppings, and whose backward calls
the library's Hellmann-Feynman adjoint (wtb_grad_energies) with the incoming
gradient as the state weights.  `SpectrumLoss` is the same map followed by the
Lorentzian absorption spectrum of the solved states and the squared misfit to
a target spectrum (wtb_grad_spectrum returns the loss with its gradient, so the
backward only scales it).  `fit` and `fit_spectrum` are small optimisation
loops on top.

Parameters of the differentiable map:
  coul    (7,)  r0, EDIEL_T, EDIEL, EDIEL_B, EDIEL_Z, W_COUL, LC
  scissor ()    scissor shift (eV)
  hop_re, hop_im (nvec, norb, norb)  real and imaginary parts of H(R), each
                entry an independent parameter of H(k) = sum_R e^{ik.R} H(R)/deg
"""
import copy

import numpy as np
import torch

from . import hip
from . import tbmodel as T


class ExcitonProblem:
    """The fixed part of the map: R vectors and degeneracies, lattice, E_F,
    k-mesh, band window and potential kind, on one GPU context."""

    def __init__(self, model, mesh, nv, nc, kind="V2DRK", ktol=0.001, nev=8, precision=hip.FP64,
                 convention=hip.CONV_REFERENCE, flags=hip.FLAG_COMPAT_F32 | hip.FLAG_GAUGE_LAPACK, device=0):
        self.model = copy.deepcopy(model)
        self.mesh = tuple(mesh)
        self.nv, self.nc, self.kind, self.ktol, self.nev = nv, nc, kind, ktol, nev
        self.precision, self.convention, self.flags = precision, convention, flags
        self.ctx = hip.Context(device)
        self.kpt = T.mp_mesh(*self.mesh, self.model.rlat)
        self.ctx.set_kpoints(self.kpt, self.mesh)
        self._state = None          # parameters of the solve held by the context
        self.report = None

    def solve(self, coul, scissor, hop):
        """Lowest nev exciton energies for the given parameters (numpy)."""
        coul = np.asarray(coul, np.float64)
        m = self.model
        m.hop = np.asarray(hop, np.complex128)
        m.scissor = float(scissor)
        self.ctx.set_tb_model(m, self.flags)
        self.ctx.hk_eig(self.nv, self.nc)
        c = dict(kind=self.kind, r0=coul[0], ediel=tuple(coul[1:4]), ez=coul[4], w=coul[5], lc=coul[6],
                 ktol=self.ktol)
        self.ctx.bse_build(c, self.precision, self.convention)
        E, _, rep = self.ctx.bse_solve(self.nev, vectors=False)
        self._state = (coul.copy(), float(scissor), m.hop.copy())
        self.report = rep
        return E

    def holds(self, coul, scissor, hop):
        s = self._state
        return (s is not None and np.array_equal(s[0], coul) and s[1] == scissor and np.array_equal(s[2], hop))

    def grad(self, w):
        return self.ctx.grad_energies(np.asarray(w, np.float64))

    def spectrum_grad(self, spec, hop=True):
        """Loss, eps2 and gradients of the spectrum misfit for the held solve;
        spec = dict(omega, target, comp, cshift, gamma)."""
        return self.ctx.grad_spectrum(spec["omega"], spec["target"], spec.get("comp", "x"),
                                      spec.get("cshift", 0.05), spec.get("gamma", 0.05), hop=hop)

    def close(self):
        self.ctx.close()


class ExcitonEnergies(torch.autograd.Function):
    """E = ExcitonEnergies.apply(problem, coul, scissor, hop_re, hop_im)."""

    @staticmethod
    def forward(ctx, problem, coul, scissor, hop_re, hop_im):
        c = coul.detach().cpu().double().numpy()
        s = float(scissor.detach().cpu())
        h = hop_re.detach().cpu().double().numpy() + 1j * hop_im.detach().cpu().double().numpy()
        E = problem.solve(c, s, h)
        ctx.problem = problem
        ctx.saved = (c, s, h)
        ctx.meta = (coul.device, hop_re.device)
        return torch.as_tensor(E, dtype=torch.float64, device=coul.device)

    @staticmethod
    def backward(ctx, gE):
        problem = ctx.problem
        c, s, h = ctx.saved
        if not problem.holds(c, s, h):
            problem.solve(c, s, h)      # another forward ran since: restore this one's state
        g = problem.grad(gE.detach().cpu().double().numpy())
        dev_c, dev_h = ctx.meta
        gc = torch.tensor([g["coulomb"][k] for k in hip.COULOMB_PARAMS], dtype=torch.float64, device=dev_c)
        gs = torch.tensor(g["scissor"], dtype=torch.float64, device=dev_c)
        gh = g["hop"]
        return (None, gc, gs, torch.as_tensor(gh.real.copy(), device=dev_h),
                torch.as_tensor(gh.imag.copy(), device=dev_h))


class SpectrumLoss(torch.autograd.Function):
    """L = SpectrumLoss.apply(problem, spec, coul, scissor, hop_re, hop_im):
    sum_j (eps2(omega_j) - target_j)^2 over the problem's nev states, spec as
    in ExcitonProblem.spectrum_grad.  The H(R) gradient is computed only when
    hop_re or hop_im requires it (the reference convention has none)."""

    @staticmethod
    def forward(ctx, problem, spec, coul, scissor, hop_re, hop_im):
        c = coul.detach().cpu().double().numpy()
        s = float(scissor.detach().cpu())
        h = hop_re.detach().cpu().double().numpy() + 1j * hop_im.detach().cpu().double().numpy()
        problem.solve(c, s, h)
        need_hop = ctx.needs_input_grad[4] or ctx.needs_input_grad[5]
        g = problem.spectrum_grad(spec, hop=need_hop)
        ctx.g = g
        ctx.meta = (coul.device, hop_re.device, need_hop)
        return torch.tensor(g["loss"], dtype=torch.float64, device=coul.device)

    @staticmethod
    def backward(ctx, gL):
        g = ctx.g
        dev_c, dev_h, need_hop = ctx.meta
        a = float(gL)
        gc = a * torch.tensor([g["coulomb"][k] for k in hip.COULOMB_PARAMS], dtype=torch.float64, device=dev_c)
        gs = a * torch.tensor(g["scissor"], dtype=torch.float64, device=dev_c)
        ghr = ghi = None
        if need_hop:
            ghr = a * torch.as_tensor(g["hop"].real.copy(), device=dev_h)
            ghi = a * torch.as_tensor(g["hop"].imag.copy(), device=dev_h)
        return None, None, gc, gs, ghr, ghi


def fit_spectrum(problem, spec, coul, scissor, hop_re, hop_im, fit_coul=(0,), fit_scissor=False, fit_hop=False,
                 steps=30, lr=1.0):
    """Least-squares fit of the absorption spectrum to spec["target"] over the
    selected parameters with L-BFGS.  Returns the fitted tensors and the loss
    history."""
    coul = coul.clone().double()
    scissor = scissor.clone().double()
    hop_re, hop_im = hop_re.clone().double(), hop_im.clone().double()
    mask = torch.zeros(7, dtype=torch.float64)
    mask[list(fit_coul)] = 1.0
    coul.requires_grad_(len(fit_coul) > 0)
    scissor.requires_grad_(fit_scissor)
    hop_re.requires_grad_(fit_hop)
    hop_im.requires_grad_(fit_hop)
    params = [p for p in (coul, scissor, hop_re, hop_im) if p.requires_grad]
    opt = torch.optim.LBFGS(params, lr=lr, max_iter=steps, line_search_fn="strong_wolfe",
                            tolerance_grad=1e-12, tolerance_change=1e-16)
    history = []

    def closure():
        opt.zero_grad()
        loss = SpectrumLoss.apply(problem, spec, coul, scissor, hop_re, hop_im)
        loss.backward()
        if coul.grad is not None:
            coul.grad.mul_(mask)
        history.append(float(loss))
        return loss

    opt.step(closure)
    return dict(coul=coul.detach(), scissor=scissor.detach(), hop_re=hop_re.detach(), hop_im=hop_im.detach(),
                history=history)


def fit(problem, target, coul, scissor, hop_re, hop_im, fit_coul=(0,), fit_scissor=False, fit_hop=False,
        steps=30, lr=1.0, weights=None):
    """Least-squares fit of the lowest exciton energies to `target` (eV) over
    the selected parameters with L-BFGS; the loss is in meV^2.  Returns the
    fitted tensors and the loss history."""
    coul = coul.clone().double()
    scissor = scissor.clone().double()
    hop_re, hop_im = hop_re.clone().double(), hop_im.clone().double()
    mask = torch.zeros(7, dtype=torch.float64)
    mask[list(fit_coul)] = 1.0
    coul.requires_grad_(len(fit_coul) > 0)
    scissor.requires_grad_(fit_scissor)
    hop_re.requires_grad_(fit_hop)
    hop_im.requires_grad_(fit_hop)
    params = [p for p in (coul, scissor, hop_re, hop_im) if p.requires_grad]
    target = torch.as_tensor(target, dtype=torch.float64)
    wts = torch.ones_like(target) if weights is None else torch.as_tensor(weights, dtype=torch.float64)
    opt = torch.optim.LBFGS(params, lr=lr, max_iter=steps, line_search_fn="strong_wolfe",
                            tolerance_grad=1e-10, tolerance_change=1e-14)
    history = []

    def closure():
        opt.zero_grad()
        E = ExcitonEnergies.apply(problem, coul, scissor, hop_re, hop_im)
        loss = (wts * (1e3 * (E[: target.numel()] - target)) ** 2).sum()
        loss.backward()
        if coul.grad is not None:
            coul.grad.mul_(mask)
        history.append(float(loss))
        return loss

    opt.step(closure)
    return dict(coul=coul.detach(), scissor=scissor.detach(), hop_re=hop_re.detach(), hop_im=hop_im.detach(),
                history=history)
