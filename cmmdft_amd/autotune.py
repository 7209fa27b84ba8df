"""Autotuner for the FFT radix sequences and tile shapes of a grid plan.

The search space per grid shape:
  * the radix sequence of each axis: every ordering of every factorisation of
    the axis length into {8, 5, 4, 3, 2} with at most one stage more than the
    shortest one;
  * the tiles: x-pass columns tb_x, y-pass columns tb_y, z-pass pencils p_z
    (density producer and final passes), and p_phi (the 6-field functional
    z pass), within the LDS and register-staging limits the library checks.
The cost is the device time of one steady-state batched Anderson iteration
(residual and Gram row, coefficients, mixing with the new density's
transform, weighted densities, pointwise derivatives, back-convolution),
measured with HIP events on the plan's stream; the search is a coordinate
descent over the seven parameters.

Results are kept per grid shape (not per batch size) in ``tuned_gfx950.json``
beside this file, and every plan of that shape picks them up when it is
created, so single-instance and batched plans of a shape share one
arithmetic order (the batched solver reproduces the plain one bit for bit).

Command line: ``python -m cmmdft_amd.autotune 128 96x96x96 ... [--out FILE]``
tunes the given shapes on the current GPU and updates the table (or FILE).
"""
from __future__ import annotations

import itertools
import json
import os
import sys
import time

import numpy as np

__all__ = ['TUNE_FILE', 'radix_candidates', 'lookup', 'apply', 'tune', 'save']

TUNE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tuned_gfx950.json')
_table = None


def _key(shape):
    return 'x'.join(str(int(n)) for n in shape)


def _load():
    global _table
    if _table is None:
        _table = {}
        if os.path.exists(TUNE_FILE):
            with open(TUNE_FILE) as fh:
                _table = json.load(fh)
    return _table


def lookup(shape):
    """Tuned settings of a grid shape, or None."""
    if os.environ.get('CMM_NO_TUNE'):
        return None
    return _load().get(_key(shape))


def apply(plan):
    """Apply the tuned settings of the plan's shape (if any); returns them."""
    cfg = lookup(plan.npoints)
    if cfg is None:
        return None
    for axis in range(3):
        plan.set_radix(axis, cfg['radix'][axis])
    plan.set_tiles(*cfg['tiles'])
    return cfg


def _factorisations(n, parts=(8, 5, 4, 3, 2)):
    out = []

    def rec(m, start, acc):
        if m == 1:
            out.append(tuple(acc))
            return
        for i in range(start, len(parts)):
            r = parts[i]
            if m % r == 0:
                rec(m // r, i, acc + [r])

    rec(n, 0, [])
    return out


def radix_candidates(n, extra_stages=1, max_stages=8):
    """Distinct radix orderings of n with at most ``extra_stages`` more stages
    than the shortest factorisation."""
    facs = [f for f in _factorisations(n) if len(f) <= max_stages]
    if not facs:
        return []
    lo = min(len(f) for f in facs)
    seqs = set()
    for f in facs:
        if len(f) <= lo + extra_stages:
            seqs.update(itertools.permutations(f))
    return sorted(seqs, key=lambda s: (len(s), s))


class _IterationBench:
    """One steady-state batched Anderson iteration (m = 5) on the plan:
    residual + Gram row, coefficients, mixing with the new density's z
    transform, and the functional chain -- the launch sequence of the batched
    solver, on history rings filled with the given density."""

    def __init__(self, plan, rho, m=5):
        import torch
        self.plan, self.m = plan, m
        nb, n = rho.shape
        self.rr = rho.reshape(1, nb, n).repeat(m + 1, 1, 1).contiguous()
        self.rs = torch.zeros((m, nb, n), dtype=torch.float64, device=rho.device)
        plan.set_instances(np.full(nb, 1e-3), np.zeros(nb))
        plan.anderson_init(np.full(nb, 0.1), np.ones(nb))

    def run(self, s):
        m, p = self.m, self.plan
        cur, rslot, new = (s - 1) % (m + 1), (s - 1) % m, s % (m + 1)
        ts = range(s - m, s)
        order_r = [t % m for t in ts]
        order_rho = [t % (m + 1) for t in ts]
        p.finish_anderson(self.rr[cur], None, False, self.rs, m, rslot, None, 1.0, 1.0)
        p.anderson_coef(m, order_r, rslot, (s - 2) % m, 1, 1.5, 0.5, 0.01, 0.8)
        p.anderson_mix(self.rr, order_rho, self.rs, order_r, self.rr[cur], self.rr[new], None, False, 1.0)
        p.derive()


def _chain_time(plan, rho, reps):
    import torch
    it = _IterationBench(plan, rho)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(plan.stream):
        plan.seed(rho, None, False, 1.0)
        plan.derive()
        it.run(6)
        ev0.record(plan.stream)
        for k in range(reps):
            it.run(7 + k)
        ev1.record(plan.stream)
    ev1.synchronize()
    return ev0.elapsed_time(ev1) / reps


def _tile_candidates(plan):
    nx, ny, nz = plan.npoints
    budget = 8 * 256
    tbx = [t for t in range(1, 33) if t * nx <= 512 and (t & (t - 1) == 0 or t == 512 // nx)]
    # y tiles: powers of two, and the widths that split the nz/2+1 columns into
    # k equal tiles (65 = 5 x 13 at 128^3: no near-empty last tile)
    nzh = nz // 2 + 1
    even = [-(-nzh // k) for k in range(1, 9)]
    # (rounded up to even: the paired y block's two rounds of ceil(tb_y / 2))
    tby = sorted({t for t in [1, 2, 4, 8, 16, 32] + even
                  if (t + (t & 1)) * ny <= budget and (t + (t & 1)) * ny * 16 <= 65536})
    pz = [p for p in (2, 4, 8, 16, 32) if (p // 2) * nz <= budget and nz * ((p // 2) | 1) * 16 <= 65536]
    if nz == 128:
        pphi = [2, 4, 6, 8]   # the compile-time z tiles of the functional pass
    else:
        pphi = [p for p in (2, 4, 6, 8, 10, 12, 16) if 3 * p * nz <= budget and nz * ((3 * p) | 1) * 16 <= 65536]
    return tbx, tby, pz, pphi


def tune(plan, functional, rho, reps=5, sweeps=2, log=None):
    """Coordinate-descent search on ``plan`` (functional set here) with the
    density ``rho`` (device, nbatch x npts); leaves the best settings applied
    and returns them as a table entry."""
    plan.set_functional(functional)
    tbx, tby, pz, pphi = _tile_candidates(plan)
    info = plan.info
    cur = {'radix': [plan.radix(a) for a in range(3)],
           'tiles': [info['tb_x'], info['tb_y'], info['p_z'], info['p_phi']]}

    def set_cfg(c):
        for a in range(3):
            plan.set_radix(a, c['radix'][a])
        plan.set_tiles(*c['tiles'])

    def cost(c):
        set_cfg(c)
        return _chain_time(plan, rho, reps)

    default_ms = best_ms = cost(cur)
    options = [('radix', a, [list(s) for s in radix_candidates(plan.npoints[a])]) for a in range(3)]
    options += [('tiles', i, [int(v) for v in vals]) for i, vals in enumerate((tbx, tby, pz, pphi))]
    for _ in range(sweeps):
        improved = False
        for kind, idx, vals in options:
            for v in vals:
                if cur[kind][idx] == v:
                    continue
                cand = {'radix': [list(r) for r in cur['radix']], 'tiles': list(cur['tiles'])}
                cand[kind][idx] = v
                t = cost(cand)
                if t < best_ms * 0.995:
                    best_ms, cur, improved = t, cand, True
                    if log:
                        log('  %s[%d] = %s -> %.4f ms' % (kind, idx, v, t))
        if not improved:
            break
    set_cfg(cur)
    return {'radix': cur['radix'], 'tiles': cur['tiles'], 'chain_ms': best_ms, 'default_chain_ms': default_ms,
            'nbatch': plan.nbatch}


def save(shape, entry, path=None):
    path = path or TUNE_FILE
    table = {}
    if os.path.exists(path):
        with open(path) as fh:
            table = json.load(fh)
    table[_key(shape)] = entry
    with open(path, 'w') as fh:
        json.dump(table, fh, indent=1, sort_keys=True)
    global _table
    if path == TUNE_FILE:
        _table = table


def _tune_shape(shape, nbatch=4):
    """Tune one shape on a representative workload: a host-like structured
    density (case-2 host at 10 bar), White-Bear II FMT + mean-field WDA."""
    import torch
    from . import units as U, workloads as W
    from ._lib import Plan
    from .eos import MFAEOS
    from .functionals import HardSphereFunctional, WDAVFunctional, build_functional
    from .system import Grid, SphericalLJGuest
    from .extpot import get_external_potential
    host = W.host_case2()
    cell = host.rvecs * (np.asarray(shape) / 128.0)[:, None]
    grid = Grid(cell, list(shape))
    guest = SphericalLJGuest('CH4', W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON)
    guest.compute_hardsphere_radius(W.TEMPERATURE)
    hs = HardSphereFunctional(grid, guest.Rhs, 'WBII')
    hs.set_temperature(W.TEMPERATURE, guest.Rhs)
    wda = WDAVFunctional(grid, guest.Rhs, MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    wda.set_temperature(W.TEMPERATURE, guest.Rhs)
    func = build_functional(hs=hs, wda=wda, kT=U.boltzmann * W.TEMPERATURE)
    pos = host.pos * (np.asarray(shape) / 128.0)[None, :]
    h2 = W.SyntheticHost(cell, pos, host.ffatype_ids, host.ffatypes, host.ff)
    V = get_external_potential(grid, host.ff, W.GUEST_SIGMA, W.GUEST_EPSILON, h2, cutoff=W.CUTOFF,
                               as_tensor=True)
    beta = 1 / (U.boltzmann * W.TEMPERATURE)
    rho1 = 2.4e-4 / U.angstrom ** 3 * torch.exp(-beta * torch.clamp(V, max=1e4 * U.kjmol))
    rho = rho1.reshape(1, -1).repeat(nbatch, 1).contiguous()
    os.environ['CMM_NO_TUNE'] = '1'
    try:
        plan = Plan(shape, cell, nbatch=nbatch)
    finally:
        del os.environ['CMM_NO_TUNE']
    t0 = time.time()
    entry = tune(plan, func, rho, log=print)
    entry['tuning_s'] = time.time() - t0
    return entry


def main(argv):
    out = None
    if '--out' in argv:
        i = argv.index('--out')
        out = argv[i + 1]
        argv = argv[:i] + argv[i + 2:]
    shapes = []
    for a in argv:
        parts = [int(v) for v in a.lower().split('x')]
        shapes.append(tuple(parts * 3) if len(parts) == 1 else tuple(parts))
    for shape in shapes:
        print('tuning', _key(shape), flush=True)
        e = _tune_shape(shape)
        print(_key(shape), json.dumps(e), flush=True)
        save(shape, e, out)


if __name__ == '__main__':
    main(sys.argv[1:])
