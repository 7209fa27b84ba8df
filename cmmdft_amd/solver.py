"""Euler-Lagrange solvers with the reference's interface (cmmdft/solver.py),
running the fused HIP pipeline.

Per iteration the device does: the functional chain of the current density
(weighted densities -> Phi derivatives -> back-convolution), then one fused
final pass that forms C1 = c_ex + V_ext, the Boltzmann update
G = beta exp(-beta C1) z and either the Picard mix or the Anderson residual
(with its Gram row), and the density producer that writes the new density
together with its z-transform.  All reductions (N, integrated |drho|, the
ideal-gas, external, FMT and WDA terms of the grand potential, Gram entries)
are computed inside those passes; the host reads one small scalar block per
iteration and runs the reference's control flow on it (convergence criteria,
Anderson's grand-potential safeguard, the hybrid Picard line search).
"""
from __future__ import annotations

import time

import numpy as np
import scipy.optimize as opt

from . import units as U
from ._device import torch_mod, as_device, like_input
from . import _lib as L
from .functionals import build_functional, configure_plan, device_parts

__all__ = ['Solver', 'Picard', 'Anderson', 'NoSolutionError', 'grand_potential']


class NoSolutionError(Exception):
    """Raised when the solver cannot find a solution (solver.py:1217-1219)."""


class _Engine:
    """Device state of one solve: plan, functional, external potential."""

    def __init__(self, grid, fener):
        torch = torch_mod()
        self.grid = grid
        self.fener = fener
        hs, wda, mfa, eps = device_parts(fener.parts)
        V = None
        for part in eps:
            V = part.potential_dev.reshape(-1) if V is None else V + part.potential_dev.reshape(-1)
        self.hs, self.wda, self.mfa = hs, wda, mfa
        self.beta = fener.beta
        self.kT = 1.0 / fener.beta
        self.lam3 = fener.wavelength ** 3
        self.func = build_functional(hs=hs, wda=wda, kT=self.kT)
        self.V = V.contiguous() if V is not None else None
        self.plan = grid.plan()
        configure_plan(self.plan, hs, wda, mfa, self.kT)
        self.dr = grid.dr
        self.npts = grid.npts
        self.dev = torch.device('cuda')

    def fex(self, s):
        """Excess energy sums outside the hard-sphere Phi (times dr): the WDA
        f_ex and the mean-field energy."""
        return s[L.S_FEX] + (s[L.S_FMFA] if self.mfa is not None else 0.0)

    def omega(self, mu, N, fid, ext, phi, fex):
        """Grand potential from the summed pieces (solver.py:103-119)."""
        dr = self.dr
        return fid * dr / self.beta - mu * N * dr + phi * dr * self.kT + fex * dr + ext * dr


def grand_potential(fener, chempot, rho):
    """Grand potential of a density with the fener's parts, on the device."""
    eng = _Engine(fener.grid, fener)
    r = as_device(rho).reshape(-1)
    eng.plan.seed(r, eng.V, False, eng.lam3)
    eng.plan.value()
    s = eng.plan.scalars()[0]
    return eng.omega(chempot, s[L.S_PN], s[L.S_PFID], s[L.S_PEXT], s[L.S_PHI], eng.fex(s))


class Solver:
    """Generic fixed-point driver (solver.py:18-345)."""
    name = 'SOLVER'

    def __init__(self, program, nsteps=500, threshold=1e-6, criterion='RIUE', a_tol=1e-6, r_tol=1e-4, min_iter=1,
                 track_history=False):
        self.grid = program.grid
        self.fener = program.fener
        self.nsteps = nsteps
        if isinstance(criterion, list):
            for crit in criterion:
                assert crit.lower() in ['riue', 'res', 'der', 'maxdrho'], \
                    'Criterion must be either RIUE, RES, DER or MAXDRHO'
            self.criterion = [c.lower() for c in criterion]
            if isinstance(threshold, list):
                assert len(criterion) == len(threshold), \
                    'If multiple criterions are given, the same amount of thresholds should be given'
                self.threshold = threshold
            else:
                self.threshold = [threshold] * len(criterion)
        else:
            assert criterion.lower() in ['riue', 'res', 'der', 'maxdrho'], \
                'Criterion must be either RIUE, RES, DER or MAXDRHO'
            if criterion.lower() == 'der':
                threshold = 1
            self.criterion = [criterion.lower()]
            self.threshold = [threshold]
        if 'der' in self.criterion:
            raise NotImplementedError("the 'DER' criterion is not implemented on the device")
        self.a_tol = a_tol
        self.r_tol = r_tol
        self.min_iter = min_iter
        self.iphase = 0
        self.log_level = 2
        self.curr_step = 0
        self.lower_density = 1e-12
        self.correction_factor = 1
        self.track_history = track_history
        self.history_header = ('Loading [au], Grand potential [Eh], Norm of residuals, RIUE, Norm of derivative, '
                               'Time per step [s], Cumulative time [s]')
        self.history = np.zeros((nsteps + 1, 7))
        self.omega_history = []

    # ---- helpers ---------------------------------------------------------
    def _initiate_solving(self, chempot):           # solver.py:90-101
        self.fugacity = np.exp(self.fener.beta * chempot) / self.fener.beta / self.fener.wavelength ** 3
        self.chempot = chempot
        self.curr_step = 0
        self.omega0 = None
        self.omega_history = []
        if self.track_history:
            self.history = np.zeros((self.nsteps + 1, 7))

    def _criteria_pass(self, N_new, iue, dmax, res2=None):
        """All criteria must pass (solver.py:167-234)."""
        self.IUE = iue * self.grid.dr
        self.RIUE = self.IUE / N_new if N_new > 0 else np.nan
        self.MAXDRHO = dmax * U.angstrom ** 3
        self.RES = np.nan
        ok = True
        for crit, th in zip(self.criterion, self.threshold):
            if crit == 'riue':
                val = self.RIUE
            elif crit == 'maxdrho':
                val = self.MAXDRHO
            else:
                self.RES = np.sqrt(res2) / np.sqrt(N_new) / np.sqrt(self.grid.npts)
                val = self.RES
            ok = ok and (val < th)
        if self.IUE == 0 and np.isnan(self.RIUE):
            ok = True
        return ok

    def _record(self, N_new, t0, tstart):
        if self.track_history and self.curr_step <= self.nsteps:
            h = self.history[self.curr_step]
            h[0] = N_new
            h[1] = self.omega0 if self.omega0 is not None else np.nan
            h[2] = self.RES
            h[3] = self.RIUE
            h[4] = np.nan
            now = time.perf_counter()
            h[5] = now - t0
            h[6] = now - tstart

    def solve(self, chempot, rho, log_level=2):     # solver.py:317-345
        self.log_level = log_level
        self.correction_factor = 1
        while self.correction_factor >= 1 / 4:
            try:
                return self._solve(chempot, rho, log_level)
            except FloatingPointError:
                self.correction_factor /= 2
                self.iphase += 1
        self.correction_factor = 1
        raise NoSolutionError('Solution not converged after %d steps at temperature %5.3f and chemical potential '
                              '%7.5f' % (self.nsteps, self.fener.temperature, chempot / U.kjmol))

    # per-solver driver
    def _solve(self, chempot, rho, log_level):
        raise NotImplementedError


class Picard(Solver):
    """Picard iteration (solver.py:347-457): 'static' fixed mixing, or 'hybrid'
    with the quadratic line search on the grand potential."""
    name = 'PICARD'

    def __init__(self, program, nsteps=500, alpha_mix=0.1, method='hybrid', **kwargs):
        super().__init__(program, nsteps, **kwargs)
        assert method in ('hybrid', 'static'), 'unknown Picard method %r' % method
        self.alpha_mix = alpha_mix
        self.method = method
        self.thresh = 1 * U.kjmol

    # ------------------------------------------------------------------
    def _alloc(self, eng, m):
        torch = torch_mod()
        n = eng.npts
        self._rho_ring = torch.empty((m + 1, n), dtype=torch.float64, device='cuda')
        self._res_ring = torch.zeros((max(m, 1), n), dtype=torch.float64, device='cuda')
        self._tmp = torch.empty(n, dtype=torch.float64, device='cuda')

    def _set_inst(self, eng, alpha):
        eng.plan.set_instances([self.fugacity], [alpha])

    def _derive_scal(self, eng):
        eng.plan.derive()

    def _density_omega(self, eng, s):
        """Omega of a density just produced (k_zproduce) and derived."""
        return eng.omega(self.chempot, s[L.S_PN], s[L.S_PFID], s[L.S_PEXT], s[L.S_PHI], eng.fex(s))

    def _value_at(self, eng, rho_slot_tensor, res_slot, alpha, clip, out):
        """Omega of rho + alpha r (trial density), value chain only."""
        self._set_inst(eng, alpha)
        eng.plan.trial(rho_slot_tensor, self._res_ring, res_slot, out, eng.V, False, eng.lam3, int(clip))
        eng.plan.value()
        s = eng.plan.scalars()[0]
        return self._density_omega(eng, s), s

    def _hybrid(self, eng, rho_cur, res_slot, n3max_cur, out):
        """update_rho_hybrid (solver.py:403-457) from rho_cur with the residual
        r = G - rho_cur stored in ``res_slot``; writes the new density to
        ``out`` and leaves its functional chain evaluated.  Returns the new
        density's scalars."""
        prev = self.omega0
        # alpha_max from max n3 of rho and of G (solver.py:150-165)
        _, sG = self._value_at(eng, rho_cur, res_slot, 1.0, False, self._tmp)
        m0, m1 = n3max_cur, sG[L.S_N3MAX]
        amax = float(np.min([abs((1 - m0) / ((m1 - m0) + 1e-16)), 1]))
        if np.isclose(amax, 0):
            aopt, lo, hi = 0.0, 0.0, 0.0
        else:
            a1 = 0.45 * amax
            o1, _ = self._value_at(eng, rho_cur, res_slot, a1, False, self._tmp)
            a2 = 0.9 * amax if o1 <= prev else 0.225 * amax
            o2, _ = self._value_at(eng, rho_cur, res_slot, a2, False, self._tmp)
            c, b, a = np.polyfit([0, a1, a2], [prev, o1, o2], 2)
            al = np.linspace(-max(a1, a2) / 4, max(a1, a2), 10000)
            om = a + b * al + c * al ** 2
            aopt = al[np.where(om == np.min(om))[0][0]]
            lo, hi = np.min(om) / U.kjmol, np.max(om) / U.kjmol
        if aopt <= 0:
            if hi - lo > self.thresh:
                f = lambda alpha: self._value_at(eng, rho_cur, res_slot, float(np.asarray(alpha).ravel()[0]),
                                                 False, self._tmp)[0]
                bounds = opt.Bounds(0.01 * amax, 0.9 * amax)
                aopt = opt.minimize(f, [self.alpha_mix * amax], bounds=bounds, method='SLSQP',
                                    options={'ftol': 1e-8}).x
            else:
                aopt = self.alpha_mix * amax
        aopt = float(np.asarray(aopt).ravel()[0])
        if aopt <= 0 or np.isclose(aopt, 0):
            aopt = self.alpha_mix * amax
        # rho_new = clip((1 - a cf) rho + a cf G)
        self._set_inst(eng, aopt * self.correction_factor)
        eng.plan.trial(rho_cur, self._res_ring, res_slot, out, eng.V, False, eng.lam3, 1)
        eng.plan.derive()
        s = eng.plan.scalars()[0]
        self.omega0 = self._density_omega(eng, s)
        return s

    def _start(self, chempot, rho):
        eng = _Engine(self.grid, self.fener)
        self._initiate_solving(chempot)
        # this solve's fugacity must be on the device before the first
        # Boltzmann update (the plan is shared between solves)
        self._set_inst(eng, 0.0)
        self.was_numpy = isinstance(rho, np.ndarray)
        r0 = as_device(rho).reshape(-1)
        return eng, r0

    def _solve(self, chempot, rho, log_level):
        eng, r0 = self._start(chempot, rho)
        self._alloc(eng, 1)
        ring = self._rho_ring
        cur, nxt = 0, 1
        ring[cur].copy_(r0)
        tstart = time.perf_counter()
        eng.plan.seed(ring[cur], eng.V, False, eng.lam3)
        eng.plan.derive()
        s = eng.plan.scalars()[0]
        self.omega0 = self._density_omega(eng, s)
        self._phi_fex = (s[L.S_PHI], eng.fex(s))
        n3max_cur = s[L.S_N3MAX]
        converged = False
        N_new = s[L.S_PN] * eng.dr
        for istep in range(self.nsteps):
            t0 = time.perf_counter()
            self.curr_step = istep + 1
            if self.method == 'static':
                self._set_inst(eng, self.alpha_mix * self.correction_factor)
                eng.plan.finish_picard(ring[cur], eng.V, False, ring[nxt], None, eng.beta, eng.lam3,
                                       1e-10 / U.angstrom ** 3)
                eng.plan.derive()
                s = eng.plan.scalars()[0]
                if s[L.S_PIC_NONFIN] > 0:
                    raise FloatingPointError('non-finite density')
                # Omega of the previous density is now complete
                self.omega_history.append(
                    eng.omega(chempot, s[L.S_NCUR], s[L.S_FID], s[L.S_EXT], self._prev_phi(s), self._prev_fex(s)))
                N_new = s[L.S_PIC_NNEW] * eng.dr
                iue, dmax = s[L.S_PIC_IUE], s[L.S_PIC_DMAX]
                self._phi_fex = (s[L.S_PHI], eng.fex(s))
            else:
                eng.plan.finish_anderson(ring[cur], eng.V, False, self._res_ring, 1, 0, None, eng.beta, eng.lam3)
                eng.plan.synchronize()
                s = self._hybrid(eng, ring[cur], 0, n3max_cur, ring[nxt])
                if s[L.S_PNONFIN] > 0:
                    raise FloatingPointError('non-finite density')
                N_new = s[L.S_PN] * eng.dr
                iue, dmax = s[L.S_PIUE], s[L.S_PDMAX]
                n3max_cur = s[L.S_N3MAX]
            res2 = self._res_norm2(eng, ring[nxt]) if 'res' in self.criterion else None
            ok = self._criteria_pass(N_new, iue, dmax, res2)
            self._record(N_new, t0, tstart)
            if ok:
                converged = True
                cur, nxt = nxt, cur
                break
            cur, nxt = nxt, cur
        out = ring[cur].reshape(self.grid.npoints).clone()
        return N_new, (out.cpu().numpy() if self.was_numpy else out), converged

    _phi_fex = (0.0, 0.0)

    def _prev_phi(self, s):
        return self._phi_fex[0]

    def _prev_fex(self, s):
        return self._phi_fex[1]

    def _res_norm2(self, eng, rho_new):
        """||G(C1(rho_new)) - rho_new||^2 for the 'RES' criterion; the
        functional chain of rho_new must be the latest one evaluated."""
        torch = torch_mod()
        scratch = torch.empty((1, eng.npts), dtype=torch.float64, device='cuda')
        eng.plan.finish_anderson(rho_new, eng.V, False, scratch, 1, 0, None, eng.beta, eng.lam3)
        s = eng.plan.scalars()[0]
        # the Gram entry is overwritten by the next step's pass; the chain state is untouched
        return s[L.S_GRAM]


class Anderson(Picard):
    """Anderson / hybrid-Anderson mixing (solver.py:459-604).  The simplex-
    constrained least-squares problem for the mixing coefficients is solved
    exactly on the device (every support's KKT system) instead of by SLSQP;
    both return its minimiser, so fixed points agree."""
    name = 'ANDERSON'

    def __init__(self, program, nsteps=500, method='hybridanderson', m=5, damping=0.3, delta=0.2, damping_max=0.8,
                 damping_min=0.01, adaptive_damping=True, damping_factors=(1.5, 0.5), **kwargs):
        super().__init__(program, nsteps, method='hybrid', **kwargs)
        assert method in ('anderson', 'hybridanderson')
        assert 1 <= m <= L.MMAX, 'history length m must be in 1..%d' % L.MMAX
        self.Anderson_method = method
        self.m = m
        self.damping = damping
        self.original_damping = damping
        self.damping_max = damping_max
        self.damping_min = damping_min
        self.adaptive_damping = adaptive_damping
        self.damping_factors = damping_factors
        self.delta = delta

    def _solve(self, chempot, rho, log_level):
        eng, r0 = self._start(chempot, rho)
        m = self.m
        self._alloc(eng, m)
        ring, rring = self._rho_ring, self._res_ring
        eng.plan.anderson_init([self.original_damping], [self.correction_factor])
        self.damping = self.original_damping
        self.And_true = False
        self.it_eps0 = np.nan
        # history bookkeeping: entries (rho slot, residual slot), oldest first
        hist = []
        free_rho = list(range(m + 1))
        cur = free_rho.pop(0)
        ring[cur].copy_(r0)
        tstart = time.perf_counter()
        eng.plan.seed(ring[cur], eng.V, False, eng.lam3)
        eng.plan.derive()
        s = eng.plan.scalars()[0]
        self.omega0 = self._density_omega(eng, s)
        n3max_cur = s[L.S_N3MAX]
        converged = False
        N_new = s[L.S_PN] * eng.dr
        nsaved = 0
        for istep in range(self.nsteps):
            t0 = time.perf_counter()
            self.curr_step = istep + 1
            prev_omega = self.omega0
            # residual of the current density into the ring (solver.py:550-552)
            rslot = nsaved % m
            eng.plan.finish_anderson(ring[cur], eng.V, False, rring, m, rslot, None, eng.beta, eng.lam3)
            # the pure method always takes the Anderson step, and the hybrid
            # one does once it has taken one (And_true stays set); then the
            # residual checks are deferred to the one host read after the step
            # (the residual pass's slots survive the mixing and derive passes)
            deferred = 'hybrid' not in self.Anderson_method.lower() or self.And_true
            if not deferred:
                sA = eng.plan.scalars()[0]
                if sA[L.S_AND_NONFIN] > 0:
                    # G of the current density is not finite: the reference falls
                    # back to the hybrid update, whose output is then non-finite too
                    raise FloatingPointError('non-finite Boltzmann factor')
            # drop the oldest entry if the history is full; keep its rho slot
            if len(hist) == m:
                old_rho, _ = hist.pop(0)
                free_rho.append(old_rho)
            hist.append((cur, rslot))
            nsaved += 1
            if not deferred:
                res_norm = np.sqrt(max(sA[L.S_GRAM + rslot], 0.0))
                N_cur = sA[L.S_NCUR] * eng.dr
                it_eps = res_norm / N_cur if N_cur != 0 else np.inf
                if self.curr_step < 3:
                    self.it_eps0 = it_eps
            cond = deferred or ((it_eps <= self.it_eps0 * self.delta) and self.curr_step > 4) or self.And_true
            nxt = free_rho.pop(0)
            used_anderson = False
            if cond:
                mk = min(len(hist), m)
                order_r = [h[1] for h in hist[-mk:]]
                order_rho = [h[0] for h in hist[-mk:]]
                slot_prev = hist[-2][1] if len(hist) >= 2 else -1
                eng.plan.anderson_coef(m, order_r, rslot, slot_prev, int(self.adaptive_damping),
                                       self.damping_factors[0], self.damping_factors[1], self.damping_min,
                                       self.damping_max)
                eng.plan.anderson_mix(ring, order_rho, rring, order_r, ring[cur], ring[nxt], eng.V, False, eng.lam3)
                eng.plan.derive()
                s = eng.plan.scalars()[0]
                if deferred:
                    if s[L.S_AND_NONFIN] > 0:
                        raise FloatingPointError('non-finite Boltzmann factor')
                    res_norm = np.sqrt(max(s[L.S_GRAM + rslot], 0.0))
                    N_cur = s[L.S_NCUR] * eng.dr
                    if self.curr_step < 3:
                        self.it_eps0 = res_norm / N_cur if N_cur != 0 else np.inf
                self.And_true = True
                used_anderson = True
                omega_new = self._density_omega(eng, s)
                bad = (not np.isfinite(s[L.S_PHI])) or (not np.isfinite(s[L.S_FEX])) or s[L.S_PNONFIN] > 0
                if bad:
                    s = self._hybrid(eng, ring[cur], rslot, n3max_cur, ring[nxt])
                    used_anderson = False
                else:
                    # the reference records Omega of the Anderson density before
                    # testing it, so a rejected step's line search starts from it
                    self.omega0 = omega_new
                    if omega_new > prev_omega * 0.8:
                        s = self._hybrid(eng, ring[cur], rslot, n3max_cur, ring[nxt])
                        used_anderson = False
            else:
                s = self._hybrid(eng, ring[cur], rslot, n3max_cur, ring[nxt])
            if s[L.S_PNONFIN] > 0:
                raise FloatingPointError('non-finite density')
            N_new = s[L.S_PN] * eng.dr
            n3max_cur = s[L.S_N3MAX]
            res2 = self._res_norm2(eng, ring[nxt]) if 'res' in self.criterion else None
            ok = self._criteria_pass(N_new, s[L.S_PIUE], s[L.S_PDMAX], res2)
            self._record(N_new, t0, tstart)
            self.last_step_anderson = used_anderson
            prev_cur = cur
            cur = nxt
            if prev_cur not in [h[0] for h in hist]:
                free_rho.append(prev_cur)
            if ok:
                converged = True
                break
        d, _ = eng.plan.anderson_state()
        self.damping = float(d[0])
        out = ring[cur].reshape(self.grid.npoints).clone()
        return N_new, (out.cpu().numpy() if self.was_numpy else out), converged
