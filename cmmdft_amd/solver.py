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
Anderson's grand-potential safeguard, the hybrid Picard line search).  The
RES and DER criteria, the solving history and the energy-tracking file need
measures of the new density's own Boltzmann update; one more final pass
(``cmm_finish_check``) reduces them when they are asked for.
"""
from __future__ import annotations

import os
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
        self.hs, self.wda, self.mfa, self.eps = hs, wda, mfa, eps
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

    def part_values(self, s, rho, ext):
        """Free energy of every part of the free energy, in the order of
        ``fener.parts`` (the columns of the tracking file): from the chain's
        sums (Phi, f_ex, mean-field energy) and the external sum ``ext`` of
        rho; several external potentials are integrated one by one."""
        out = []
        for part in self.fener.parts:
            if part.name == 'HardSphere':
                out.append(s[L.S_PHI] * self.dr * self.kT)
            elif part.name == 'WDA-V':
                out.append(s[L.S_FEX] * self.dr)
            elif part.name == 'MFA':
                out.append(s[L.S_FMFA] * self.dr)
            elif len(self.eps) == 1:
                out.append(ext * self.dr)
            else:
                out.append(part.value(rho.reshape(self.grid.npoints)))
        return out


def grand_potential(fener, chempot, rho):
    """Grand potential of a density with the fener's parts, on the device."""
    eng = _Engine(fener.grid, fener)
    r = as_device(rho).reshape(-1)
    eng.plan.seed(r, eng.V, False, eng.lam3)
    eng.plan.value()
    s = eng.plan.scalars()[0]
    return eng.omega(chempot, s[L.S_PN], s[L.S_PFID], s[L.S_PEXT], s[L.S_PHI], eng.fex(s))


class Solver:
    """Generic fixed-point driver (solver.py:18-345).  Besides the
    reference's criteria 'RIUE', 'RES' and 'DER' it accepts 'MAXDRHO'
    (max |rho_new - rho| in angstrom^-3)."""
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
        self.a_tol = a_tol
        self.r_tol = r_tol
        self.min_iter = min_iter
        self.iphase = 0
        self.log_level = 2
        self.curr_step = 0
        self.lower_density = 1e-12
        self.correction_factor = 1
        self.fugacity = None
        self.track_history = track_history
        self.history_header = ('Loading [au], Grand potential [Eh], Norm of residuals, RIUE, Norm of derivative, '
                               'Time per step [s], Cumulative time [s]')
        self.history = np.zeros((nsteps + 1, 7))
        self.omega_history = []
        self._eng = None

    # ---- reference helpers ------------------------------------------------
    def _initiate_solving(self, chempot):           # solver.py:90-101
        self.fugacity = np.exp(self.fener.beta * chempot) / self.fener.beta / self.fener.wavelength ** 3
        self.chempot = chempot
        self.curr_step = 0
        self.omega0 = None
        self.omega_history = []
        self._eng = None
        if self.track_history:
            self.history = np.zeros((self.nsteps + 1, 7))

    def get_new_rho(self, C1, fugacity):            # solver.py:128-129
        """beta exp(-beta C1) fugacity, on the device (numpy in, numpy out)."""
        torch = torch_mod()
        c = as_device(C1)
        out = torch.empty_like(c)
        L.boltzmann(c, self.fener.beta, fugacity, out)
        return like_input(out, C1)

    def _get_C1(self, rho, krho=None):              # solver.py:137-144
        """Sum of the parts' derivatives at rho, on the device."""
        torch = torch_mod()
        eng = _Engine(self.grid, self.fener)
        r = as_device(rho).reshape(-1)
        eng.plan.seed(r, eng.V, False, eng.lam3)
        eng.plan.derive()
        c1 = torch.empty_like(r)
        eng.plan.finish_c1(r, eng.V, False, c1, eng.lam3)
        return like_input(c1.reshape(self.grid.npoints), rho)

    def _get_Omega(self, rho, krho=None):           # solver.py:103-119
        self.omega0 = grand_potential(self.fener, self.chempot, rho)
        return self.omega0

    # ---- convergence and bookkeeping --------------------------------------
    def _need_check(self, produced):
        """Whether the new density needs the extra measuring pass: for the RES
        and DER criteria, the history, and (for the static Picard update,
        whose final pass produces the density without its sums) tracking."""
        return ('res' in self.criterion or 'der' in self.criterion or self.track_history
                or (self.fener.fn_tracking is not None and not produced))

    def _measure(self, eng, rho, s, produced):
        """Measures of a density whose functional chain was just derived:
        N, ideal and external sums (from the producer pass, or the check
        pass), Phi / f_ex / mean-field sums (from the chain) and, when the
        check pass ran, |G - rho|^2 and the DER sums."""
        m = dict(phi=s[L.S_PHI], fex=eng.fex(s), s=s, chk=None)
        if produced:
            m.update(N=s[L.S_PN], fid=s[L.S_PFID], ext=s[L.S_PEXT])
        if self._need_check(produced):
            eng.plan.finish_check(rho, eng.V, False, eng.beta, eng.lam3, self.a_tol, self.r_tol)
            c = eng.plan.scalars()[0]
            m['chk'] = dict(nonfin=c[L.S_CHK_NONFIN], res2=c[L.S_CHK_RES2], der2=c[L.S_CHK_DER2],
                            dern2=c[L.S_CHK_DERN2])
            if not produced:
                m.update(N=c[L.S_NCUR], fid=c[L.S_FID], ext=c[L.S_EXT])
        return m

    def _criteria_pass(self, N_new, iue, dmax, meas):
        """All criteria must pass (solver.py:167-234)."""
        npts = self.grid.npts
        self.IUE = iue * self.grid.dr
        self.RIUE = self.IUE / N_new if N_new > 0 else np.nan
        self.MAXDRHO = dmax * U.angstrom ** 3
        self.RES = np.nan
        self.DER = np.nan
        chk = meas['chk'] if meas is not None else None
        ok = True
        for crit, th in zip(self.criterion, self.threshold):
            if crit == 'riue':
                val = self.RIUE
            elif crit == 'maxdrho':
                val = self.MAXDRHO
            elif crit == 'res':
                self.RES = np.sqrt(chk['res2']) / np.sqrt(N_new) / np.sqrt(npts)
                val = self.RES
            else:
                self.DER = np.sqrt(chk['der2']) / np.sqrt(npts)
                val = self.DER
            ok = ok and (val < th)
        if self.IUE == 0 and np.isnan(self.RIUE):
            ok = True
        return ok

    def _omega_of(self, eng, meas):
        return eng.omega(self.chempot, meas['N'], meas['fid'], meas['ext'], meas['phi'], meas['fex'])

    def _track_size(self):
        """Length of the energy-tracking file (None without one)."""
        fn = self.fener.fn_tracking
        return os.path.getsize(fn) if fn is not None and os.path.exists(fn) else None

    def _untrack(self, size):
        """Drops the tracking lines written after the file had ``size`` bytes."""
        fn = self.fener.fn_tracking
        if fn is not None and size is not None and os.path.exists(fn):
            with open(fn, 'r+') as f:
                f.truncate(size)

    def _track(self, eng, meas, rho):
        """One line of the energy-tracking file for the density just measured
        (Solver._get_Omega / _track_energy, solver.py:103-126)."""
        fn = self.fener.fn_tracking
        if fn is None:
            return
        dr, mu = self.grid.dr, self.chempot
        N = meas['N'] * dr
        Fid = meas['fid'] * dr / eng.beta
        G = Fid - mu * N
        line = '%6i\t%4i\t%.6e\t%.6e\t% .6e' % (self.iphase, self.curr_step, N, (-mu * N), Fid)
        for F in eng.part_values(meas['s'], rho, meas['ext']):
            G += F
            line += '\t% .6e' % F
        line += '\t% .6e' % G
        with open(fn, 'a') as f:
            f.write(line + '\n')

    def _record(self, N_new, meas, t0, tstart):
        """A row of the solving history (solver.py:207-223)."""
        if not self.track_history or self.curr_step > self.nsteps:
            return
        npts = self.grid.npts
        h = self.history[self.curr_step]
        chk = meas['chk']
        h[0] = N_new
        h[1] = self.omega0 if self.omega0 is not None else np.nan
        h[2] = self.RES if not np.isnan(self.RES) else np.sqrt(chk['res2']) / N_new / np.sqrt(npts)
        h[3] = self.RIUE
        h[4] = np.sqrt(chk['dern2']) / np.sqrt(npts)
        now = time.perf_counter()
        h[5] = now - t0
        h[6] = now - tstart

    def _record_start(self, eng, rho, meas):
        """Row 0 of the history and line 0 of the tracking file
        (solver.py:261-278)."""
        self._track(eng, meas, rho)
        if self.track_history:
            chk = meas['chk']
            npts = self.grid.npts
            self.history[0] = [meas['N'] * self.grid.dr, self.omega0, np.sqrt(chk['res2']), np.nan,
                               np.sqrt(chk['dern2']) / np.sqrt(npts), np.nan, np.nan]

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
    _nring = 1

    def _alloc(self, eng, m, nrho=None):
        torch = torch_mod()
        n = eng.npts
        self._rho_ring = torch.empty((nrho or m + 1, n), dtype=torch.float64, device='cuda')
        self._res_ring = torch.zeros((max(m, 1), n), dtype=torch.float64, device='cuda')
        self._tmp = torch.empty(n, dtype=torch.float64, device='cuda')

    def _set_inst(self, eng, alpha):
        eng.plan.set_instances([self.fugacity], [alpha])

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
        self._initiate_solving(chempot)
        eng = _Engine(self.grid, self.fener)
        # this solve's fugacity must be on the device before the first
        # Boltzmann update (the plan is shared between solves)
        self._set_inst(eng, 0.0)
        self.was_numpy = isinstance(rho, np.ndarray)
        r0 = as_device(rho).reshape(-1)
        return eng, r0

    def _begin(self, eng, r0):
        """Device state of a solve from r0: rings, the seeded and derived
        first density, its grand potential (solver.py:256-278)."""
        self._eng = eng
        self._alloc(eng, self._nring)
        self._cur, self._nxt = 0, 1
        self._rho_ring[0].copy_(r0)
        eng.plan.seed(self._rho_ring[0], eng.V, False, eng.lam3)
        eng.plan.derive()
        s = eng.plan.scalars()[0]
        self.omega0 = self._density_omega(eng, s)
        self._phi_fex = (s[L.S_PHI], eng.fex(s))
        self._n3max_cur = s[L.S_N3MAX]
        meas = self._measure(eng, self._rho_ring[0], s, True)
        self._record_start(eng, self._rho_ring[0], meas)
        return s

    def _step(self):
        """update_rho from the current density (solver.py:392-457): the new
        density goes to ring slot self._nxt with its chain derived."""
        eng = self._eng
        ring, cur, nxt = self._rho_ring, self._cur, self._nxt
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
                eng.omega(self.chempot, s[L.S_NCUR], s[L.S_FID], s[L.S_EXT], self._phi_fex[0], self._phi_fex[1]))
            self._phi_fex = (s[L.S_PHI], eng.fex(s))
            return dict(s=s, N=s[L.S_PIC_NNEW] * eng.dr, iue=s[L.S_PIC_IUE], dmax=s[L.S_PIC_DMAX], produced=False)
        eng.plan.finish_anderson(ring[cur], eng.V, False, self._res_ring, 1, 0, None, eng.beta, eng.lam3)
        sA = eng.plan.scalars()[0]
        if sA[L.S_AND_NONFIN] > 0:
            # G of the current density is not finite: so is the line search's
            # new density (solver.py:288-290)
            raise FloatingPointError('non-finite Boltzmann factor')
        s = self._hybrid(eng, ring[cur], 0, self._n3max_cur, ring[nxt])
        if s[L.S_PNONFIN] > 0:
            raise FloatingPointError('non-finite density')
        return dict(s=s, N=s[L.S_PN] * eng.dr, iue=s[L.S_PIUE], dmax=s[L.S_PDMAX], produced=True)

    def _advance(self):
        self._cur, self._nxt = self._nxt, self._cur

    def _finish_step(self, info, t0, tstart):
        """Criteria, tracking and history of the density a step produced;
        makes it the current density.  Returns whether all criteria pass."""
        eng = self._eng
        rho_new = self._rho_ring[self._nxt]
        s = info['s']
        meas = self._measure(eng, rho_new, s, info['produced'])
        if not info['produced'] and 'N' in meas:
            self.omega0 = self._omega_of(eng, meas)
        self._n3max_cur = s[L.S_N3MAX]
        ok = self._criteria_pass(info['N'], info['iue'], info['dmax'], meas)
        self._track(eng, meas, rho_new)
        self._record(info['N'], meas, t0, tstart)
        self._advance()
        return ok

    def _confirm(self):
        """None when a density that meets the criteria is accepted; else the
        step taken again (see Anderson)."""
        return None

    def _solve(self, chempot, rho, log_level):
        eng, r0 = self._start(chempot, rho)
        tstart = time.perf_counter()
        s = self._begin(eng, r0)
        converged = False
        N_new = s[L.S_PN] * eng.dr
        istep = 0
        while istep < self.nsteps:
            t0 = time.perf_counter()
            self.curr_step = istep + 1
            info = self._step()
            if info.get('redo'):
                # the previous step was taken again (Anderson: its density
                # had a non-finite Boltzmann factor)
                istep -= 1
                self.curr_step = istep + 1
            N_new = info['N']
            ok = self._finish_step(info, t0, tstart)
            if ok:
                info = self._confirm()
                if info is None:
                    converged = True
                    break
                N_new = info['N']
                if self._finish_step(info, t0, tstart):
                    converged = True
                    break
            istep += 1
        out = self._rho_ring[self._cur].reshape(self.grid.npoints).clone()
        return N_new, (out.cpu().numpy() if self.was_numpy else out), converged

    def update_rho(self, rho, krho=None, C1=None):
        """One update of the density (Picard.update_rho_static /
        update_rho_hybrid, Anderson.update_rho; solver.py:392-457, 546-579)
        after ``_initiate_solving(chempot)``, with the step number in
        ``curr_step`` as the reference's loop sets it.  Returns (rho_new,
        krho_new, C1_new) in the reference conventions (numpy in, numpy out).
        The device recomputes C1 from rho: ``krho`` and ``C1`` are accepted
        for the signature and must belong to rho, as in the reference's loop.
        Consecutive calls that pass the previous call's rho_new continue the
        device state (the Anderson history among it); any other rho starts a
        new one."""
        torch = torch_mod()
        if self.fugacity is None:
            raise AssertionError('update_rho needs _initiate_solving(chempot) first')
        r = as_device(rho).reshape(-1)
        if self._eng is None or not torch.equal(r, self._rho_ring[self._cur]):
            eng = _Engine(self.grid, self.fener)
            self._set_inst(eng, 0.0)
            self._begin(eng, r)
        eng = self._eng
        self.curr_step = max(self.curr_step, 1)
        info = self._step()
        self._n3max_cur = info['s'][L.S_N3MAX]
        self._advance()
        rho_new = self._rho_ring[self._cur]
        c1 = torch.empty_like(rho_new)
        eng.plan.finish_c1(rho_new, eng.V, False, c1, eng.lam3)
        rho_out = rho_new.reshape(self.grid.npoints).clone()
        krho_new = self.grid.fft(rho_out)
        return (like_input(rho_out, rho), like_input(krho_new, rho) if isinstance(rho, np.ndarray) else krho_new,
                like_input(c1.reshape(self.grid.npoints), rho))

    _phi_fex = (0.0, 0.0)


class Anderson(Picard):
    """Anderson / hybrid-Anderson mixing (solver.py:459-604).  The simplex-
    constrained least-squares problem for the mixing coefficients is solved
    exactly on the device (every support's KKT system) instead of by SLSQP;
    both return its minimiser, so fixed points agree.  With
    ``track_history`` the coefficients and damping of every Anderson step
    are kept in ``anderson_log`` as (step, coefficients, damping)."""
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
        self.anderson_log = []
        # steps whose Anderson density is treated as having a non-finite
        # Boltzmann factor (a host-side switch for tests of the fallback;
        # empty in use), and the steps taken again as hybrid steps
        self._force_nonfinite = ()
        self.redone_steps = []
        self._rb = None

    def _initiate_solving(self, chempot):           # solver.py:515-525
        super()._initiate_solving(chempot)
        self.And_true = False
        self.it_eps0 = np.nan
        self.damping = self.original_damping
        self.anderson_log = []

    def _begin(self, eng, r0):
        m = self.m
        self._eng = eng
        # m + 1 density slots hold the history and the new density; one more
        # keeps a step's starting density while the next step runs, in case
        # that step has to be taken again (_redo)
        self._alloc(eng, m, m + 2)
        eng.plan.anderson_init([self.original_damping], [self.correction_factor])
        self.damping = self.original_damping
        self.And_true = False
        self.it_eps0 = np.nan
        self.last_step_anderson = False
        self.redone_steps = []
        self._rb = None
        # history bookkeeping: entries (rho slot, residual slot), oldest first
        self._hist = []
        self._free = list(range(m + 2))
        self._cur = self._free.pop(0)
        self._nsaved = 0
        ring = self._rho_ring
        ring[self._cur].copy_(r0)
        eng.plan.seed(ring[self._cur], eng.V, False, eng.lam3)
        eng.plan.derive()
        s = eng.plan.scalars()[0]
        self.omega0 = self._density_omega(eng, s)
        self._n3max_cur = s[L.S_N3MAX]
        meas = self._measure(eng, ring[self._cur], s, True)
        self._record_start(eng, ring[self._cur], meas)
        return s

    def _step(self):
        """Anderson.update_rho (solver.py:546-579) from the current density."""
        eng, m = self._eng, self.m
        ring, rring = self._rho_ring, self._res_ring
        hist, free = self._hist, self._free
        cur = self._cur
        prev_omega = self.omega0
        # residual of the current density into the ring (solver.py:550-552)
        rslot = self._nsaved % m
        eng.plan.finish_anderson(ring[cur], eng.V, False, rring, m, rslot, None, eng.beta, eng.lam3)
        # the pure method always takes the Anderson step, and the hybrid one
        # does once it has taken one (And_true stays set); then the residual
        # checks are deferred to the one host read after the step (the
        # residual pass's slots survive the mixing and derive passes)
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
            free.append(old_rho)
        hist.append((cur, rslot))
        self._nsaved += 1
        if not deferred:
            res_norm = np.sqrt(max(sA[L.S_GRAM + rslot], 0.0))
            N_cur = sA[L.S_NCUR] * eng.dr
            it_eps = res_norm / N_cur if N_cur != 0 else np.inf
            if self.curr_step < 3:
                self.it_eps0 = it_eps
        cond = deferred or ((it_eps <= self.it_eps0 * self.delta) and self.curr_step > 4) or self.And_true
        # this step's bookkeeping, to take it again as a hybrid step should
        # its Anderson density turn out to have a non-finite Boltzmann factor
        rb = dict(cur=cur, rslot=rslot, omega=prev_omega, n3max=self._n3max_cur, hist=list(hist),
                  free=list(free), nsaved=self._nsaved, track=self._track_size())
        nxt = free.pop(0)
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
                forced = self.last_step_anderson and (self.curr_step - 1) in self._force_nonfinite
                if s[L.S_AND_NONFIN] > 0 or forced:
                    # G of the current density is not finite.  If an
                    # Anderson step produced it, the reference would have
                    # taken a hybrid step there instead (solver.py:567-568)
                    if self.last_step_anderson and self._rb is not None:
                        return self._redo(self.curr_step - 1)
                    raise FloatingPointError('non-finite Boltzmann factor')
                res_norm = np.sqrt(max(s[L.S_GRAM + rslot], 0.0))
                N_cur = s[L.S_NCUR] * eng.dr
                if self.curr_step < 3:
                    self.it_eps0 = res_norm / N_cur if N_cur != 0 else np.inf
            rb['damping'] = s[L.S_ADAMP]
            self._rb = rb
            if self.track_history:
                d, a = eng.plan.anderson_state()
                self.anderson_log.append((self.curr_step, a[0, :mk].copy(), float(d[0])))
            self.And_true = True
            used_anderson = True
            omega_new = self._density_omega(eng, s)
            bad = (not np.isfinite(s[L.S_PHI])) or (not np.isfinite(s[L.S_FEX])) or s[L.S_PNONFIN] > 0
            if bad:
                s = self._hybrid(eng, ring[cur], rslot, self._n3max_cur, ring[nxt])
                used_anderson = False
            else:
                # the reference records Omega of the Anderson density before
                # testing it, so a rejected step's line search starts from it
                self.omega0 = omega_new
                if omega_new > prev_omega * 0.8:
                    s = self._hybrid(eng, ring[cur], rslot, self._n3max_cur, ring[nxt])
                    used_anderson = False
        else:
            # a step without mixing still saves its residual into the history
            # (solver.py:550-552): the Gram ring gets this step's row
            eng.plan.anderson_store_row(m, rslot)
            s = self._hybrid(eng, ring[cur], rslot, self._n3max_cur, ring[nxt])
        if s[L.S_PNONFIN] > 0:
            raise FloatingPointError('non-finite density')
        self.last_step_anderson = used_anderson
        self._nxt = nxt
        return dict(s=s, N=s[L.S_PN] * eng.dr, iue=s[L.S_PIUE], dmax=s[L.S_PDMAX], produced=True)

    def _redo(self, step):
        """Takes the last step (``step``) again as a hybrid step from its starting
        density: its Anderson density has a non-finite Boltzmann factor, so
        the reference would have taken update_rho_hybrid there
        (solver.py:561-568).  The history, the free ring slots, the grand
        potential and max n3 of the starting density and the tracking file
        return to that step; the damping keeps that step's update (the
        reference adapts it before testing the new density,
        solver.py:596-597) and drops the later one.  Returns the step's
        info, flagged as a redo."""
        rb = self._rb
        self._rb = None
        eng, m = self._eng, self.m
        ring, rring = self._rho_ring, self._res_ring
        self._hist, self._free, self._nsaved = list(rb['hist']), list(rb['free']), rb['nsaved']
        self._cur = cur = rb['cur']
        self.omega0 = rb['omega']
        self._n3max_cur = rb['n3max']
        eng.plan.anderson_set_damping([rb['damping']])
        self._untrack(rb['track'])
        if m == 1:
            # the rejected density's residual took the only ring slot: the
            # starting density's residual (and Gram row) again
            eng.plan.seed(ring[cur], eng.V, False, eng.lam3)
            eng.plan.derive()
            eng.plan.finish_anderson(ring[cur], eng.V, False, rring, m, rb['rslot'], None, eng.beta, eng.lam3)
        nxt = self._free.pop(0)
        s = self._hybrid(eng, ring[cur], rb['rslot'], self._n3max_cur, ring[nxt])
        if s[L.S_PNONFIN] > 0:
            raise FloatingPointError('non-finite density')
        self.last_step_anderson = False
        self._nxt = nxt
        self.redone_steps.append(step)
        return dict(s=s, N=s[L.S_PN] * eng.dr, iue=s[L.S_PIUE], dmax=s[L.S_PDMAX], produced=True, redo=True)

    def _confirm(self):
        """A density that meets the criteria and came from an Anderson step is
        accepted only if its Boltzmann factor is finite (one check pass);
        else the step is taken again as a hybrid step (_redo), as the
        reference's update_rho would have (solver.py:567-568)."""
        if not self.last_step_anderson or self._rb is None:
            return None
        eng = self._eng
        eng.plan.finish_check(self._rho_ring[self._cur], eng.V, False, eng.beta, eng.lam3, self.a_tol, self.r_tol)
        c = eng.plan.scalars()[0]
        if c[L.S_CHK_NONFIN] == 0 and self.curr_step not in self._force_nonfinite:
            return None
        return self._redo(self.curr_step)

    def _advance(self):
        prev_cur = self._cur
        self._cur = self._nxt
        if prev_cur not in [h[0] for h in self._hist]:
            self._free.append(prev_cur)

    def _solve(self, chempot, rho, log_level):
        out = super()._solve(chempot, rho, log_level)
        d, _ = self._eng.plan.anderson_state()
        self.damping = float(d[0])
        return out
