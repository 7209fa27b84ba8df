"""Batched Euler-Lagrange solves: many state points of one host and grid
(the chemical potentials of an isotherm) iterated together on the device.

The reference loops over state points one by one
(examples/methane_ZIF8.ipynb, the cascade_solver loop).  Here every kernel
of the pipeline runs over the whole batch at once (one grid dimension of every
launch is the instance), each instance carries its own Anderson history,
damping, grand potential and convergence state on the device, and a device
kernel retires converged instances from an activity mask that every later
kernel reads.  The instances advance in lockstep, so their history-ring slots
are common; after the first m steps the slot pattern is periodic with period
m (m + 1).  That period is captured once, as m HIP graphs of m + 1
iterations each, which are replayed in turn until every instance has
finished.  The host reads the status block between graph launches, so a
batch that has finished runs at most m + 1 further (fully masked)
iterations: with one graph of the whole period, up to m (m + 1) masked
iterations ran after the last instance converged (measured at 128^3, 32
instances: 18 masked iterations, ~10 ms of a 110 ms isotherm).

Instances are grouped in chunks (``chunk`` per group, each group its own
plan and history rings); within an iteration every group runs its whole
pipeline before the next group starts.  By default every instance goes into
one group, capped by device memory (``mem_cap`` bytes for spectra and
history rings, 7 half spectra plus 2m + 1 real fields per instance, about
0.3 GB per instance at 128^3 with m = 5): one launch per pass over the whole
batch amortises launch and tail effects best.  Measured at 128^3 on MI355X,
a 16-instance group ran 1.14x faster than 2-instance groups, whose working
set would fit the 256 MB Infinity Cache, so cache residency does not pay for
the extra launches there.

Per instance the arithmetic is that of the plain ``Anderson(method='anderson')``
solver (same kernels, same tiles, same reduction order), so a state point that
converges without the grand-potential safeguard firing gives the same result
bit for bit.  Instances whose safeguard fires (the reference would take a
hybrid Picard step there) or that stop on non-finite values are re-solved
with the plain solver.
"""
from __future__ import annotations

import numpy as np

from . import units as U
from ._device import torch_mod, as_device
from . import _lib as L
from .functionals import configure_plan, device_parts
from .solver import Anderson, NoSolutionError

__all__ = ['BatchedAnderson', 'UnevenBatch', 'footprint_bytes']

STATUS = {0: 'running', 1: 'converged', 2: 'step limit', 3: 'fallback', 4: 'non-finite'}


def _nfields(fener):
    hs = fener.part_dict.get('HardSphere') if fener is not None else None
    return 12 if hs is not None and 't' in hs.version else 6


def footprint_bytes(npoints, ninst, m=5, nfields=6):
    """Device bytes of a batched solve: per instance nfields + 1 half spectra
    of the plan workspace (nfields = 6, or 12 for the tensor FMT variant), the
    density (m + 1) and residual (m) history rings, plus the density output
    and external potential of the grid."""
    nx, ny, nz = (int(n) for n in npoints)
    npts, nhalf = nx * ny * nz, nx * ny * (nz // 2 + 1)
    return ninst * ((nfields + 1) * 16 * nhalf + (2 * m + 2) * 8 * npts) + 8 * npts


class BatchedAnderson:
    def __init__(self, program, chempots, m=5, damping=0.1, threshold=1e-8, nsteps=500, use_graph=True,
                 adaptive_damping=True, damping_max=0.8, damping_min=0.01, damping_factors=(1.5, 0.5),
                 resolve_fallbacks=True, chunk=None, mem_cap=64 << 30):
        assert 1 <= m <= L.MMAX
        self.program = program
        self.grid = program.grid
        self.fener = program.fener
        self.chempots = np.asarray(chempots, dtype=np.float64).ravel()
        self.m = m
        self.damping = damping
        self.threshold = threshold
        self.nsteps = nsteps
        self.use_graph = use_graph
        self.adaptive = adaptive_damping
        self.dmax, self.dmin = damping_max, damping_min
        self.factors = damping_factors
        self.resolve_fallbacks = resolve_fallbacks
        if chunk is None:
            nx, ny, nz = self.grid.npoints
            per_inst = (_nfields(self.fener) + 1) * 16 * nx * ny * (nz // 2 + 1) + (2 * m + 1) * 8 * nx * ny * nz
            chunk = max(1, min(len(self.chempots), int(mem_cap // per_inst)))
        self.chunk = int(chunk)
        self._groups = None
        self._graphs = None     # the period's graphs, replayed in turn
        self._gi = 0            # the next graph to replay
        self._graph_key = None
        self._graph_V = None

    # ------------------------------------------------------------------
    def _setup(self):
        torch = torch_mod()
        fener, grid = self.fener, self.grid
        B = len(self.chempots)
        hs, wda, mfa, eps = device_parts(fener.parts)
        V = None
        for part in eps:
            V = part.potential_dev.reshape(-1) if V is None else V + part.potential_dev.reshape(-1)
        self.beta = fener.beta
        self.lam3 = fener.wavelength ** 3
        self.V = V.contiguous() if V is not None else None
        sizes = [min(self.chunk, B - i) for i in range(0, B, self.chunk)]
        if self._groups is None or [g['n'] for g in self._groups] != sizes:
            from ._lib import Plan
            self.close()
            self.stream = torch.cuda.Stream()
            n = grid.npts
            self._groups = []
            first = 0
            for nb in sizes:
                plan = Plan(grid.npoints, grid.cell.rvecs, nbatch=nb)
                plan.bind_stream(self.stream)
                self._groups.append(dict(
                    n=nb, first=first, plan=plan,
                    rho_ring=torch.empty((self.m + 1, nb, n), dtype=torch.float64, device='cuda'),
                    res_ring=torch.zeros((self.m, nb, n), dtype=torch.float64, device='cuda')))
                first += nb
        for g in self._groups:
            mu = self.chempots[g['first']:g['first'] + g['n']]
            g['mu'] = mu
            configure_plan(g['plan'], hs, wda, mfa, 1.0 / self.beta)
            g['plan'].set_instances(np.exp(self.beta * mu) / self.beta / self.lam3, np.zeros(g['n']))
        # A captured graph holds by value what its launches were given: the
        # external-potential pointer, beta, Lambda^3, the functional and the
        # mean-field table pointer.  Re-capture when any of them changed since
        # (a new temperature, another potential array, other parts).
        p0 = self._groups[0]['plan']
        key = (0 if self.V is None else self.V.data_ptr(), self.beta, self.lam3, bytes(p0._func),
               0 if p0._mfa is None else p0._mfa.data_ptr())
        if self._graphs is not None and key != self._graph_key:
            self._destroy_graphs()
        self._graph_key = key

    def _iteration(self, s):
        """Enqueue batched Anderson step s (1-based) for every group."""
        m = self.m
        cur, rslot, new = (s - 1) % (m + 1), (s - 1) % m, s % (m + 1)
        mk = min(s, m)
        ts = range(s - mk, s)
        order_r = [t % m for t in ts]
        order_rho = [t % (m + 1) for t in ts]
        slot_prev = (s - 2) % m if s >= 2 else -1
        for g in self._groups:
            p, rr, rs = g['plan'], g['rho_ring'], g['res_ring']
            p.finish_anderson(rr[cur], self.V, False, rs, m, rslot, None, self.beta, self.lam3)
            p.anderson_coef(m, order_r, rslot, slot_prev, int(self.adaptive), self.factors[0], self.factors[1],
                            self.dmin, self.dmax)
            p.anderson_mix(rr, order_rho, rs, order_r, rr[cur], rr[new], self.V, False, self.lam3)
            p.derive()
            p.batch_check(new)

    def _state(self):
        sts = [g['plan'].batch_state() for g in self._groups]
        return {k: np.concatenate([st[k] for st in sts]) for k in sts[0]}

    def start(self, rho0=None):
        """Upload the state points, seed every instance and run the first m
        (warm-up) iterations; capture the periodic part into a HIP graph.
        Work is enqueued on this solver's stream."""
        torch = torch_mod()
        self._setup()
        B, m = len(self.chempots), self.m
        self.stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self.stream):
            self._Vd = self.V if self.V is not None else torch.zeros(self.grid.npts, dtype=torch.float64,
                                                                     device='cuda')
            self._rho0 = rho0
            for g in self._groups:
                p, rr = g['plan'], g['rho_ring']
                if rho0 is None:
                    # ideal-gas initial densities (program.py:296-305), one per state point
                    mu = torch.from_numpy(g['mu']).to('cuda')[:, None]
                    rr[0].copy_(torch.exp(self.beta * (mu - self._Vd[None, :])) / self.lam3)
                else:
                    rr[0].copy_(as_device(rho0).reshape(B, -1)[g['first']:g['first'] + g['n']])
                g['res_ring'].zero_()
                p.anderson_init(np.full(g['n'], self.damping), np.ones(g['n']))
                p.batch_reset(g['mu'], self.beta, self.lam3, self.grid.dr, self.threshold, self.nsteps)
                p.seed(rr[0], self.V, False, self.lam3)
                p.derive()
                p.batch_check(0, init=True)
            s = 0
            # warm-up: the first m steps have growing history lengths
            while s < m and s < self.nsteps:
                s += 1
                self._iteration(s)
            self._s = s
            # no read of the status block here (it would wait for this
            # stream): solvers started one after another (UnevenBatch) then
            # run their warm-ups concurrently.  The period is captured even
            # when every instance has already finished; replays over
            # finished instances are fully masked and change nothing.
            self._st = None
            self._graphed = bool(self.use_graph and s < self.nsteps)
            if self._graphed and self._graphs is None:
                p0 = self._groups[0]['plan']
                seg = m + 1
                self._graphs = []
                for j in range(m):
                    p0.graph_begin()
                    try:
                        for k in range(j * seg, (j + 1) * seg):
                            self._iteration(m + 1 + k)
                    finally:
                        self._graphs.append(p0.graph_end())
                # the graphs read this potential array: keep it alive with them
                self._graph_V = self.V
            self._gi = 0
        return self

    def running(self):
        """True while some instance is active and the step limit is not reached
        (as of the last poll)."""
        if self._st is None:   # not polled since start()
            return self._s < self.nsteps
        return bool(self._st['active'].any()) and self._s < self.nsteps

    def advance(self):
        """Enqueue the next block of iterations (the next of the period's
        graphs, m + 1 iterations, or up to 8 eager iterations); does not
        wait."""
        torch = torch_mod()
        if self._graphed:
            self._groups[0]['plan'].graph_launch(self._graphs[self._gi])
            self._gi = (self._gi + 1) % len(self._graphs)
            self._s += self.m + 1
        else:
            with torch.cuda.stream(self.stream):
                for _ in range(8):
                    if self._s >= self.nsteps:
                        break
                    self._s += 1
                    self._iteration(self._s)

    def poll(self):
        """Read the device status block (waits for this solver's stream)."""
        self._st = self._state()
        return self.running()

    def finish(self):
        """Gather the converged densities and scalars; re-solve fallback,
        step-limit and non-finite instances with the plain solver (an instance
        whose re-solve raises NoSolutionError keeps its status, with N and rho
        NaN).  Returns dict(N, rho (B, nx, ny, nz) device tensor, converged,
        iters, status, omega, resolved)."""
        torch = torch_mod()
        B = len(self.chempots)
        if self._st is None:
            self._st = self._state()
        st = self._st
        with torch.cuda.stream(self.stream):
            scal = np.concatenate([g['plan'].scalars() for g in self._groups])
            rho = torch.empty((B,) + tuple(self.grid.npoints), dtype=torch.float64, device='cuda')
            for g in self._groups:
                for j in range(g['n']):
                    i = g['first'] + j
                    rho[i].copy_(g['rho_ring'][st['final_slot'][i], j].reshape(self.grid.npoints))
        self.stream.synchronize()
        for g in self._groups:
            g['plan'].batch_release()
        N = scal[:, L.S_PN] * self.grid.dr
        status = st['status'].copy()
        iters = st['iters'].copy()
        redo = [i for i in range(B) if status[i] in (2, 3, 4)] if self.resolve_fallbacks else []
        for i in redo:
            solver = Anderson(self.program, nsteps=self.nsteps, method='anderson', m=self.m, damping=self.damping,
                              criterion='RIUE', threshold=self.threshold, adaptive_damping=self.adaptive,
                              damping_max=self.dmax, damping_min=self.dmin, damping_factors=self.factors)
            mu = self.chempots[i]
            r0 = torch.exp(self.beta * (mu - self._Vd.reshape(self.grid.npoints))) / self.lam3 \
                if self._rho0 is None else as_device(self._rho0)[i]
            try:
                Ni, ri, ci = solver.solve(mu, r0)
            except NoSolutionError:
                # the plain solver fails too (solver.py:335-345): the state
                # point keeps its status, without a density
                rho[i].fill_(float('nan'))
                N[i] = float('nan')
                iters[i] = solver.curr_step
                continue
            rho[i].copy_(ri.reshape(self.grid.npoints))
            N[i] = Ni
            iters[i] = solver.curr_step
            status[i] = 1 if ci else 2
        return dict(N=N, rho=rho, converged=status == 1, iters=iters, status=status, omega=st['omega'],
                    resolved=redo)

    def solve(self, rho0=None):
        """Returns dict(N, rho (B, nx, ny, nz) device tensor, converged, iters,
        status, omega, resolved)."""
        self.start(rho0)
        while self.running():
            self.advance()
            self.poll()
        return self.finish()

    def footprint(self):
        """Device bytes this solver's groups hold (spectra, rings, scalars)."""
        return footprint_bytes(self.grid.npoints, len(self.chempots), self.m, _nfields(self.fener))

    def _destroy_graphs(self):
        if self._graphs is not None and self._groups:
            for gr in self._graphs:
                self._groups[0]['plan'].graph_destroy(gr)
        self._graphs = None
        self._graph_V = None

    def close(self):
        self._destroy_graphs()


class UnevenBatch:
    """State points of many frameworks of different grid sizes, solved
    concurrently on one GPU.

    ``items`` is a list of (program, chempots): one program per framework (its
    grid, host and functional).  Each framework's state points form one
    BatchedAnderson (its own plan, stream and HIP graph); frameworks are
    admitted by device footprint -- largest first, while the admitted total
    stays within ``mem_cap`` bytes -- and run concurrently: every round
    launches one graph replay per running framework on its own stream, then
    polls them, retires the finished ones and admits more.
    """

    def __init__(self, items, m=5, damping=0.1, threshold=1e-8, nsteps=500, use_graph=True, mem_cap=None,
                 chunk=None):
        self.items = list(items)
        self.m, self.damping, self.threshold, self.nsteps = m, damping, threshold, nsteps
        self.use_graph = use_graph
        self.chunk = chunk
        if mem_cap is None:
            torch = torch_mod()
            free, _total = torch.cuda.mem_get_info()
            mem_cap = int(0.6 * free)
        self.mem_cap = int(mem_cap)
        self.stats = {}

    def footprints(self):
        return [footprint_bytes(p.grid.npoints, len(mu), self.m, _nfields(p.fener)) for p, mu in self.items]

    def solve(self):
        """Returns a list (one per item) of BatchedAnderson result dicts."""
        fps = self.footprints()
        order = sorted(range(len(self.items)), key=lambda i: -fps[i])
        if order and fps[order[0]] > self.mem_cap:
            raise ValueError('framework %d needs %d bytes, above the memory cap %d'
                             % (order[0], fps[order[0]], self.mem_cap))
        pending = list(order)
        running = {}
        results = [None] * len(self.items)
        used = peak_used = peak_running = rounds = 0
        admitted = []
        while pending or running:
            # admission: largest pending framework that fits, repeatedly
            progress = True
            while pending and progress:
                progress = False
                for j, i in enumerate(pending):
                    if used + fps[i] <= self.mem_cap:
                        prog, mus = self.items[i]
                        b = BatchedAnderson(prog, mus, m=self.m, damping=self.damping, threshold=self.threshold,
                                            nsteps=self.nsteps, use_graph=self.use_graph,
                                            chunk=self.chunk or len(mus))
                        b.start()
                        running[i] = b
                        used += fps[i]
                        admitted.append(i)
                        pending.pop(j)
                        progress = True
                        break
            peak_used = max(peak_used, used)
            peak_running = max(peak_running, len(running))
            rounds += 1
            for i, b in list(running.items()):
                if b.running():
                    b.advance()
            for i, b in list(running.items()):
                if not b.poll():
                    results[i] = b.finish()
                    b.close()
                    del running[i]
                    used -= fps[i]
        self.stats = dict(admission_order=admitted, peak_bytes=peak_used, peak_concurrent=peak_running,
                          rounds=rounds, mem_cap=self.mem_cap)
        return results
