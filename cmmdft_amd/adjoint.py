"""Derivatives of the adsorbed amount N = integral of the converged density
with respect to the chemical potential and the host LJ parameters.

The reference has no derivatives of the Euler-Lagrange solution; its fixed
point is rho = Lambda^-3 exp(beta (mu - C1(rho))) (Solver.get_new_rho,
cmmdft/solver.py:128-129) with C1 = c_ex(rho) + V_ext.  At that point the
residual R(rho) = kT ln(Lambda^3 rho) + c_ex(rho) + V - mu vanishes, and by the
implicit function theorem

    dN/dp = -lambda^T dR/dp,   J lambda = dN/drho = dr,   J = diag(kT/rho) + M,

with M = dc_ex/drho (symmetric).  So dN/dmu = sum lambda and
dN/dtheta = -sum lambda dV/dtheta.  The device solves J lambda = dr as
(I + S M S) y = S dr, S = sqrt(beta rho), by conjugate gradients; M v is the
functional chain with the pointwise stage differentiated by dual numbers, and
sum lambda dV/dtheta is the backward of the LJ grid sum.  One adjoint solve
gives the derivatives with respect to mu and every host parameter at once.

``uptake`` wraps forward solve + adjoint in a ``torch.autograd.Function``.
"""
from __future__ import annotations

import numpy as np

from ._device import torch_mod, as_device
from . import _lib as L
from .solver import _Engine

__all__ = ['AdjointSolver', 'BatchedAdjoint', 'uptake', 'set_host_parameters', 'host_parameters']


def _check_symmetric(fener):
    """The conjugate-gradient adjoint needs a symmetric Hessian of the
    excess free energy.  The reference's mean-field convolution keeps one
    factor of its transform phase exp(i pi scalprod), which displaces the
    pair potential, so its operator is not symmetric: not supported here."""
    if 'MFA' in fener.part_names:
        raise NotImplementedError('adjoint with the mean-field (MFA) part: its convolution operator is not '
                                  'symmetric (the reference phase displaces the pair potential)')


class AdjointSolver:
    """Linear response of a converged density of ``program`` (its grid and
    free energy)."""

    def __init__(self, program, tol=1e-10, maxiter=1000):
        self.program = program
        self.grid = program.grid
        self.fener = program.fener
        _check_symmetric(self.fener)
        self.tol = tol
        self.maxiter = maxiter
        self.iters = 0
        self.residual = None

    def _engine(self):
        return _Engine(self.grid, self.fener)

    def _prepare(self, eng, rho):
        torch = torch_mod()
        r = as_device(rho).reshape(-1).contiguous()
        nfld = torch.empty(eng.plan.nfields * self.grid.npts, dtype=torch.float64, device=r.device)
        eng.plan.seed(r, None, False, eng.lam3)
        eng.plan.weighted_densities(nfld)
        return r, nfld

    def hessian_vector(self, rho, v):
        """M v = (d c_ex / d rho) v at ``rho`` (device tensor, npoints shape)."""
        torch = torch_mod()
        eng = self._engine()
        r, nfld = self._prepare(eng, rho)
        vv = as_device(v).reshape(-1).contiguous()
        out = torch.empty_like(vv)
        eng.plan.seed(vv, None, False, eng.lam3)
        eng.plan.hessian_apply(nfld)
        eng.plan.finish_c1(vv, None, False, out, eng.lam3)
        return out.reshape(self.grid.npoints)

    def solve(self, rho):
        """lambda with J lambda = dr at the converged ``rho``; returns
        (lambda device tensor, dN/dmu)."""
        torch = torch_mod()
        eng = self._engine()
        plan = eng.plan
        r, nfld = self._prepare(eng, rho)
        n = self.grid.npts
        y, res, p, q = (torch.empty(n, dtype=torch.float64, device=r.device) for _ in range(4))
        plan.adjoint_seed(r, eng.beta, eng.dr, 1, 0, y, res, p)
        s = plan.scalars()[0]
        bb = s[L.S_ADJ_RR]
        self.iters = 0
        self.residual = 0.0
        if bb > 0.0:
            for k in range(self.maxiter):
                par = k & 1
                plan.hessian_apply(nfld)
                plan.adjoint_final(r, eng.beta, p, q)
                plan.adjoint_update(par, y, res, p, q)
                s = plan.scalars()[0]
                self.iters = k + 1
                if not s[L.S_ADJ_PQ] > 0.0:
                    raise FloatingPointError('adjoint operator is not positive definite at this density '
                                             '(p.Jp = %r): the state is not a stable minimum' % s[L.S_ADJ_PQ])
                self.residual = np.sqrt(s[L.S_ADJ_RR + (par ^ 1)] / bb)
                if not np.isfinite(self.residual):
                    raise FloatingPointError('non-finite adjoint residual')
                if self.residual <= self.tol:
                    break
                plan.adjoint_seed(r, eng.beta, eng.dr, 0, par ^ 1, y, res, p)
            else:
                raise RuntimeError('adjoint CG did not converge in %d iterations (residual %.3e)'
                                   % (self.maxiter, self.residual))
        lam = torch.empty(n, dtype=torch.float64, device=r.device)
        plan.adjoint_lambda(r, eng.beta, y, lam)
        s = plan.scalars()[0]
        return lam.reshape(self.grid.npoints), float(s[L.S_ADJ_LAM])

    def host_gradients(self, lam):
        """(ntypes, 2) array of dN/dsigma_t, dN/depsilon_t for the host
        force-field types (order of ``host.mol.ffatypes``), given lambda."""
        return _host_grad(self.fener, self.grid, as_device(lam).reshape(-1).contiguous())

    def gradients(self, rho):
        """dict(dN_dmu, dN_dsigma (ntypes,), dN_depsilon (ntypes,), iters)."""
        lam, dmu = self.solve(rho)
        out = dict(dN_dmu=dmu, iters=self.iters, residual=self.residual, lam=lam)
        ep = self.fener.part_dict.get('ExtPot')
        if ep is not None and getattr(ep, 'generated', None) is not None:
            g = self.host_gradients(lam)
            out['dN_dsigma'] = g[:, 0]
            out['dN_depsilon'] = g[:, 1]
        return out


def _host_grad(fener, grid, lam):
    """-(sum over the weight fields of lam) . dV/dtheta for the host types:
    (ntypes, 2) array of d/dsigma_t, d/depsilon_t."""
    torch = torch_mod()
    ep = fener.part_dict.get('ExtPot')
    if ep is None or getattr(ep, 'generated', None) is None:
        raise ValueError('host-parameter gradients need an external potential generated from the host '
                         '(add_external_potential without potential= or load_fn=)')
    host, guest = ep.host, ep.guest
    from .extpot import host_atom_table
    tab = as_device(host_atom_table(host.ff, host.mol))
    types = torch.from_numpy(np.asarray(host.mol.ffatype_ids, dtype=np.int32)).to(tab.device)
    g = L.lj_grid_grad(grid.npoints, np.asarray(grid.cell.rvecs, dtype=np.float64), tab, types, tab.shape[0],
                       len(host.mol.ffatypes), guest.sigma, guest.epsilon, ep.cutoff, int(ep.generated['images']),
                       ep.limit_potential, lam)
    return -g.cpu().numpy()


class BatchedAdjoint:
    """Adjoint solves for a batch of converged densities of one program (the
    state points of one framework), in lockstep: one plan of B instances, the
    CG scalars per instance; iterates until every instance's relative residual
    is below ``tol``.  ``gradients`` also returns the derivatives of the total
    uptake sum_i N_i with respect to the host type parameters, from one LJ
    backward pass over sum_i lambda_i."""

    def __init__(self, program, tol=1e-10, maxiter=1000):
        self.program = program
        self.grid = program.grid
        self.fener = program.fener
        _check_symmetric(self.fener)
        self.tol = tol
        self.maxiter = maxiter
        self.iters = 0
        self.residual = None

    def solve(self, rho):
        """rho: (B, nx, ny, nz) device tensor.  Returns (lambda (B, npts),
        dN_i/dmu_i (B,))."""
        torch = torch_mod()
        from ._lib import Plan
        from .functionals import build_functional
        hs = self.fener.part_dict.get('HardSphere')
        wda = self.fener.part_dict.get('WDA-V')
        r = as_device(rho).reshape(rho.shape[0], -1).contiguous()
        B, n = r.shape
        beta = self.fener.beta
        lam3 = self.fener.wavelength ** 3
        plan = Plan(self.grid.npoints, self.grid.cell.rvecs, nbatch=B)
        plan.set_functional(build_functional(hs=hs, wda=wda, kT=1.0 / beta))
        nfld = torch.empty((B, plan.nfields, n), dtype=torch.float64, device=r.device)
        plan.seed(r, None, False, lam3)
        plan.weighted_densities(nfld)
        y, res, p, q = (torch.empty((B, n), dtype=torch.float64, device=r.device) for _ in range(4))
        plan.adjoint_seed(r, beta, self.grid.dr, 1, 0, y, res, p)
        bb = plan.scalars()[:, L.S_ADJ_RR].copy()
        bb[bb == 0.0] = 1.0
        self.iters = 0
        for k in range(self.maxiter):
            par = k & 1
            plan.hessian_apply(nfld)
            plan.adjoint_final(r, beta, p, q)
            plan.adjoint_update(par, y, res, p, q)
            s = plan.scalars()
            self.iters = k + 1
            rel = np.sqrt(s[:, L.S_ADJ_RR + (par ^ 1)] / bb)
            if not np.all(np.isfinite(rel)):
                raise FloatingPointError('non-finite adjoint residual')
            self.residual = float(rel.max())
            if self.residual <= self.tol:
                break
            plan.adjoint_seed(r, beta, self.grid.dr, 0, par ^ 1, y, res, p)
        else:
            raise RuntimeError('batched adjoint CG did not converge in %d iterations (residual %.3e)'
                               % (self.maxiter, self.residual))
        lam = torch.empty((B, n), dtype=torch.float64, device=r.device)
        plan.adjoint_lambda(r, beta, y, lam)
        return lam, plan.scalars()[:, L.S_ADJ_LAM].copy()

    def gradients(self, rho):
        """dict(dN_dmu (B,), dNtot_dsigma (ntypes,), dNtot_depsilon (ntypes,),
        iters, residual)."""
        lam, dmu = self.solve(rho)
        g = _host_grad(self.fener, self.grid, lam)
        return dict(dN_dmu=dmu, dNtot_dsigma=g[:, 0], dNtot_depsilon=g[:, 1], iters=self.iters,
                    residual=self.residual)


def host_parameters(program):
    """(sigma, epsilon) arrays of the host force-field types, in the order of
    ``host.mol.ffatypes``."""
    host = program.system.host
    sig = np.array([host.ff[t][0] for t in host.mol.ffatypes], dtype=np.float64)
    eps = np.array([host.ff[t][1] for t in host.mol.ffatypes], dtype=np.float64)
    return sig, eps


def set_host_parameters(program, sigma, epsilon):
    """Set the host type parameters and regenerate the external potential
    on the device (clamped as at setup)."""
    host = program.system.host
    for t, name in enumerate(host.mol.ffatypes):
        host.ff[name] = (float(sigma[t]), float(epsilon[t]))
    ep = program.fener.part_dict['ExtPot']
    ep.generate_potential(images=bool(ep.generated['images']) if ep.generated else False)
    ep.clamp()


def uptake(program, mu, sigma, epsilon, solver=None, adjoint_tol=1e-10):
    """N(mu, sigma_t, epsilon_t) as a differentiable torch scalar.

    mu: 0-d float64 tensor; sigma, epsilon: (ntypes,) float64 tensors of the
    host type parameters (atomic units).  The forward pass sets the host
    parameters, regenerates V_ext, and solves with ``solver`` (default: the
    program's); the backward pass runs one adjoint solve."""
    return _Uptake.apply(mu, sigma, epsilon, program, solver, adjoint_tol)


def _make_uptake():
    torch = torch_mod()

    class Uptake(torch.autograd.Function):
        @staticmethod
        def forward(ctx, mu, sigma, epsilon, program, solver, adjoint_tol):
            set_host_parameters(program, sigma.detach().cpu().numpy(), epsilon.detach().cpu().numpy())
            if solver is not None:
                program.set_solver(solver)
            N, rho, conv = program.solve(float(mu))
            if not conv:
                raise RuntimeError('forward solve did not converge')
            ctx.program = program
            ctx.adjoint_tol = adjoint_tol
            ctx.rho = as_device(rho)
            return torch.tensor(float(N), dtype=torch.float64)

        @staticmethod
        def backward(ctx, gN):
            adj = AdjointSolver(ctx.program, tol=ctx.adjoint_tol)
            g = adj.gradients(ctx.rho)
            s = float(gN)
            gm = torch.tensor(s * g['dN_dmu'], dtype=torch.float64)
            gs = torch.from_numpy(s * g['dN_dsigma'])
            ge = torch.from_numpy(s * g['dN_depsilon'])
            return gm, gs, ge, None, None, None

    return Uptake


class _LazyUptake:
    _cls = None

    def apply(self, *args):
        if _LazyUptake._cls is None:
            _LazyUptake._cls = _make_uptake()
        return _LazyUptake._cls.apply(*args)


_Uptake = _LazyUptake()
