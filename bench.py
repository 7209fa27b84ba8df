#!/usr/bin/env python
"""Benchmark: converged isotherm points per second at 128^3 on one MI355X.

Workload: the 128^3 host (cubic a = 25.6 A, 276 LJ atoms of 4 types, seeded),
united-atom methane at 300 K, White-Bear II FMT + mean-field WDA + LJ external
potential, Anderson mixing (m = 5, damping 0.1, RIUE < 1e-8), over the
isotherm of 32 pressures log-spaced 1e-2..1e2 bar (mu from the MBWR EOS).
A "step" is one full isotherm: all 32 state points solved to convergence from
their ideal-gas initial densities.  ``value`` is the batched rate (every
state point iterated together, HIP-graph replay); the plain rate (the drop-in
solver, one state point after the other) is reported beside it.  The external
potential is computed before the timed region and reported separately.

Prints one JSON line (rank 0).  Default: N = 1, 2 timed isotherms per mode,
1 warm-up.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'isotherm points/s at 128^3, batched vs plain; HBM GB/s; adjoint/fwd time'
PEAK_HBM_GBPS = 8000.0
NSIDE = 128


def build_program(shape):
    from cmmdft_amd import units as U, workloads as W
    from cmmdft_amd.program import Program
    from cmmdft_amd.system import NanoporousHost, SphericalLJGuest
    from cmmdft_amd.eos import MFAEOS
    host = W.host_case2()
    prog = Program(write_files=False)
    prog.set_system(NanoporousHost('case2', host, host.ff),
                    SphericalLJGuest('CH4_UA', W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    prog.set_grid(npoints=list(shape))
    prog.init_free_energy(W.TEMPERATURE)
    prog.fener.add_hard_sphere('WBII')
    prog.fener.add_wdav(MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    return host, prog


def chempots(n=32):
    from cmmdft_amd import workloads as W
    from cmmdft_amd.eos import ModifiedBenedictWebbRubinEOS
    eos = ModifiedBenedictWebbRubinEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON)
    return np.array([eos.calculate_mu(W.TEMPERATURE, p) for p in W.pressures_isotherm(n)])


PASSES = (0, 1, 2, 3)


def chain_fields(plan):
    """Fields the z pass of the six-field functional transforms: five (n2,
    n3, n_v2 x and y, rho_bar; the z component of n_v2 is i kz times n3 and
    is formed in the z pass), four when the WDA is linear and applied in k
    space (rho_bar is never formed).  The paired y passes move them (the y
    component formed from n3, or folded into n3's derivative, there); the x
    passes one fewer."""
    return 4 if plan.info['wda_kspace'] else 5


def pass_names(nyf):
    return {0: 'k_zphi', 1: 'k_xweights', 2: 'k_xback', 3: 'k_ypass(%d fields)' % nyf}


def pass_bytes(nx, ny, nz, nyf):
    """Algorithmic HBM bytes per launch of each functional pass (one
    instance) with ``nyf`` fields travelling through the z pass: every
    complex half-spectrum element read or written once.  The y component of
    n_v2 is formed by the paired y pass from n3 (and its derivative folded
    into n3's there), so the x passes move one field fewer: k_xweights reads
    rho^ and writes nyf - 1 fields; k_xback reads nyf - 1 (with a linear WDA
    three of them and rho^) and writes one; the y pass timed (the back
    chain's) reads nyf fields and writes nyf - 1."""
    nh = nx * ny * (nz // 2 + 1) * 16
    return {0: 2 * nyf * nh, 1: nyf * nh, 2: 5 * nh, 3: (2 * nyf - 1) * nh}


def measured_traffic(kernel, n, nbatch=1):
    """HBM bytes per launch of ``kernel`` at n^3 (a launch covering ``nbatch``
    instances) from the committed FETCH_SIZE/WRITE_SIZE summaries
    (profiles/functional_passes_128cube*.md), or None when no measurement
    for that grid and batch is recorded."""
    for name in ('functional_passes_128cube.md', 'functional_passes_128cube_batch32.md'):
        fn = os.path.join(ROOT, 'profiles', name)
        if not os.path.exists(fn):
            continue
        for line in open(fn):
            w = line.split()
            if len(w) in (4, 5) and w[0] == 'traffic_bytes_per_launch' and w[1] == kernel.replace(' ', '_') \
                    and int(w[2]) == n and (int(w[3][1:]) if len(w) == 5 else 1) == nbatch:
                return int(w[-1])
    return None


def zpass_input(plan, rho, V, lam3):
    """A callable that puts the functional z pass's input in place: seeds the
    plan with rho (every instance) and evaluates the forward chain
    (plan.value(), whose z pass stores nothing), leaving the weighted
    densities in the workspace."""
    def prep():
        plan.seed(rho, V, False, lam3)
        plan.value()
    return prep


def time_passes(plan, prep, reps=20):
    """Average duration of each pass, HIP events on the plan's stream.  The
    z pass overwrites its input with its output, and its pointwise stage
    branches on the values (series and closed forms), so each of its timed
    launches follows an untimed prep() (zpass_input), which puts weighted
    densities in place; the other passes do the same work whatever the
    values."""
    import torch
    out = {}
    for which in PASSES:
        plan.launch_pass(which)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        if which == 0:
            total = 0.0
            for _ in range(reps):
                prep()
                e0.record(plan.stream)
                plan.launch_pass(which)
                e1.record(plan.stream)
                e1.synchronize()
                total += e0.elapsed_time(e1)
            out[which] = total / reps
            continue
        e0.record(plan.stream)
        for _ in range(reps):
            plan.launch_pass(which)
        e1.record(plan.stream)
        e1.synchronize()
        out[which] = e0.elapsed_time(e1) / reps
    return out


def host_cpu_model():
    """The host CPU's model name with its family / model numbers and the
    logical CPUs (/proc/cpuinfo), for the CPU baseline (a virtualised host
    may report only the brand in the model name)."""
    info, ncpu = {}, 0
    try:
        for line in open('/proc/cpuinfo'):
            k, _, v = line.partition(':')
            k = k.strip()
            if k == 'processor':
                ncpu += 1
            if k in ('model name', 'cpu family', 'model', 'stepping') and k not in info:
                info[k] = v.strip()
    except OSError:
        return None
    if 'model name' not in info:
        return None
    return '%s (family %s, model %s, stepping %s; %d logical CPUs)' % (
        info['model name'], info.get('cpu family', '?'), info.get('model', '?'), info.get('stepping', '?'), ncpu)


def cpu_baseline(V, shape, mu, gpu_iters_per_point, nsample, cores=1):
    """The oracle (NumPy/SciPy restatement of the reference path; scipy.fft
    with ``cores`` workers, BLAS/OpenMP pools limited to as many) timed for
    ``nsample`` Anderson iterations at the same grid and potential;
    extrapolated to points/s with the GPU's mean iteration count per
    point."""
    import scipy.fft as sfft
    from threadpoolctl import threadpool_limits
    from cmmdft_amd import units as U, workloads as W
    from oracle import cdft_oracle as O
    host = W.host_case2()
    T = W.TEMPERATURE
    with threadpool_limits(limits=cores), sfft.set_workers(cores):
        og = O.Grid(host.rvecs, shape)
        R = O.bh_radius(W.GUEST_SIGMA, W.GUEST_EPSILON, T)
        lam = O.thermal_wavelength(W.GUEST_MASS, T)
        beta = 1 / (U.boltzmann * T)
        parts = [O.HardSphere(og, R, 'WBII', T),
                 O.WDAV(og, R, O.MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON), T), O.ExtPot(og, V)]
        stamps = []

        class Timed(O.Anderson):
            def converged(self, *a, **k):
                stamps.append(time.perf_counter())
                return super().converged(*a, **k)

        s = Timed(og, parts, beta, lam, nsteps=nsample + 1, method='anderson', m=5, damping=0.1,
                  criterion='RIUE', threshold=1e-30)
        rho0 = O.initial_density(og, mu, beta, lam, V=parts[2].potential)
        s.solve(mu, rho0)
    per_iter = float(np.median(np.diff(stamps))) if len(stamps) > 1 else float('nan')
    pts = 1.0 / (per_iter * gpu_iters_per_point)
    return {'value': pts, 'unit': 'points/s', 'cores': cores, 'kind': 'port', 'host_cpu': host_cpu_model(),
            'sample': ('oracle Anderson (SLSQP mixing, scipy.fft %d worker%s) at %dx%dx%d, %d iterations timed, '
                       'median %.3f s/iteration, extrapolated with the GPU mean of %.1f iterations/point'
                       % ((cores, '' if cores == 1 else 's') + shape + (nsample, per_iter, gpu_iters_per_point))),
            'seconds_per_iteration': per_iter}


def adjoint_leg(shape, reps=3):
    """Forward solve vs adjoint gradients at one state point: the case-4 host
    (case-1 cell a = 24 A, 64 atoms in 4 LJ types) at 10 bar on the bench
    grid.  Forward: Anderson (m = 5, damping 0.1, RIUE < 1e-8) from the
    ideal-gas density.  Adjoint: the linear-response CG solve (relative
    residual 1e-10) plus dN/dmu and the 8 host-parameter gradients.  Medians
    over ``reps`` after one warm-up."""
    import torch
    from cmmdft_amd import units as U, workloads as W
    from cmmdft_amd.program import Program
    from cmmdft_amd.system import NanoporousHost, SphericalLJGuest
    from cmmdft_amd.eos import MFAEOS, ModifiedBenedictWebbRubinEOS
    from cmmdft_amd.solver import Anderson
    from cmmdft_amd.adjoint import AdjointSolver
    host = W.host_case4()
    prog = Program(write_files=False)
    prog.set_system(NanoporousHost('case4', host, host.ff),
                    SphericalLJGuest('CH4_UA', W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    prog.set_grid(npoints=list(shape))
    prog.init_free_energy(W.TEMPERATURE)
    prog.fener.add_hard_sphere('WBII')
    prog.fener.add_wdav(MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    prog.fener.add_external_potential(cutoff=W.CUTOFF)
    mu = ModifiedBenedictWebbRubinEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON).calculate_mu(W.TEMPERATURE,
                                                                                                  10 * U.bar)
    solver = Anderson(prog, nsteps=500, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-8)
    adj = AdjointSolver(prog, tol=1e-10)
    tf, ta = [], []
    for k in range(reps + 1):
        rho0 = prog._set_initial_density(chempot=mu, temperature=W.TEMPERATURE)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        N, rho, conv = solver.solve(mu, rho0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g = adj.gradients(rho)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if k > 0:
            tf.append(t1 - t0)
            ta.append(t2 - t1)
    fwd, bwd = float(np.median(tf)), float(np.median(ta))
    return {'workload': 'case-4 host (a=24 A, 64 atoms, 4 LJ types) at %dx%dx%d, 10 bar, WBII + MFA-WDA' % shape,
            'forward_ms': fwd * 1e3, 'adjoint_ms': bwd * 1e3, 'adjoint_vs_forward': bwd / fwd,
            'forward_iterations': int(solver.curr_step), 'cg_iterations': int(adj.iters),
            'cg_relative_residual': float(adj.residual), 'converged': bool(conv), 'N': float(N),
            'dN_dmu': g['dN_dmu'], 'dN_dsigma': [float(v) for v in g['dN_dsigma']],
            'dN_depsilon': [float(v) for v in g['dN_depsilon']]}


def effpot_leg(shape, degree=9):
    """Orientationally averaged external potential of a rigid CO2 (3 LJ
    sites) in the bench host (case-2, a = 25.6 A, 276 atoms) on the bench
    grid at 300 K, with ``degree`` Lebedev directions x ``degree`` angles
    about the molecular axis (cmm_effpot, one launch).  Median of 3 after a
    warm-up; the kernel alone, timed with HIP events on the current stream."""
    import torch
    from cmmdft_amd import units as U, workloads as W
    from cmmdft_amd.system import Grid
    from cmmdft_amd.extpot import get_external_potential_dict, generate_effective_potential
    from cmmdft_amd.rotations import combined_rotations
    host = W.host_case2()
    mol, ff = W.guest_co2()
    grid = Grid(host.rvecs, list(shape))
    gens = get_external_potential_dict(host.ff, ff, host, cutoff=W.CUTOFF)
    beta = 1 / (U.boltzmann * W.TEMPERATURE)
    ts = []
    for k in range(4):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        V = generate_effective_potential(mol, grid, beta, gens, degree=degree, as_tensor=True)
        e1.record()
        torch.cuda.synchronize()
        if k > 0:
            ts.append(e0.elapsed_time(e1))
    ms = float(np.median(ts))
    # the interpolated form: 0.15 A derivative tables of the cell per site type
    # (the reference's interpolate=True recipe), then the interpolated average
    from cmmdft_amd.extpot import get_external_potential_derivatives, Interpolator
    h = 0.15 * U.angstrom
    egrid = Grid(host.rvecs, spacing=h)
    tt, ti = [], []
    for k in range(3):
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        igens = {}
        for atom, (sf, ef) in ff.items():
            t = get_external_potential_derivatives(egrid, host.ff, sf, ef, host, egrid.spacings, cutoff=W.CUTOFF,
                                                   as_tensor=True)
            igens[atom] = Interpolator(t, grid.points[0, 0, 0, :3], [h, h, h]).tricubic
            igens[atom].table()
        e1.record()
        Vi = generate_effective_potential(mol, grid, beta, igens, degree=degree, as_tensor=True)
        e2.record()
        torch.cuda.synchronize()
        if k > 0:
            tt.append(e0.elapsed_time(e1))
            ti.append(e1.elapsed_time(e2))
        del igens, t
    from cmmdft_amd.extpot import _effpot_setup, _merge_orientations
    rot, wts = combined_rotations(degree)
    nrot = len(wts)
    # orientations the kernel evaluates: those with distinct site offsets
    # (a linear guest's rotations about its axis coincide)
    ndist = len(_merge_orientations(_effpot_setup(mol, gens, rot)[4], wts)[1])
    npts = int(np.prod(shape))
    sites = npts * ndist * mol.natom
    # in-cut-off pairs per site: host atoms in the cut-off sphere (uniform estimate)
    vol = abs(np.linalg.det(host.rvecs))
    pairs = sites * host.natom * (4 / 3 * np.pi * W.CUTOFF ** 3) / vol
    return {'workload': 'rigid CO2 (3 sites) in the case-2 host at %dx%dx%d, 300 K, %d orientations '
                        '(Lebedev degree %d x %d about the axis), r_c 12 A' % (shape + (nrot, degree, degree)),
            'ms': ms, 'orientations': nrot, 'distinct_orientations': ndist,
            'site_evaluations_per_s': sites / (ms * 1e-3),
            'in_cutoff_pairs_per_s_est': pairs / (ms * 1e-3),
            'fp64_gflops_est': pairs * 30 / (ms * 1e-3) / 1e9,
            'min_kjmol': float(V.min()) / U.kjmol,
            'interpolated': {'tables_ms': float(np.median(tt)), 'average_ms': float(np.median(ti)),
                             'table_grid': list(egrid.npoints),
                             'site_evaluations_per_s': sites / (float(np.median(ti)) * 1e-3),
                             'min_kjmol': float(Vi.min()) / U.kjmol}}


def uneven_leg(nframe=16, npress=8):
    """The uneven batch: synthetic frameworks k = 0..nframe-1 (orthorhombic
    edges 18-40 A, 0.25 A grids 72..160 per axis, 100-600 atoms of 4 LJ types,
    periodic images), each at npress pressures 0.1..50 bar -- all solves
    batched on one GPU (UnevenBatch, admission by device footprint), then the
    batched adjoint of every framework: dN_i/dmu_i and the gradient of the
    framework's total uptake with respect to its 8 host parameters.  External
    potentials are generated before the timed region: vext_s is the set-up of
    the programs (image-sum LJ grids, clamps, mu(p)), frameworks_s the
    host-side random placement of the synthetic atoms."""
    import torch
    from cmmdft_amd import workloads as W
    from cmmdft_amd.batch import UnevenBatch
    from cmmdft_amd.adjoint import BatchedAdjoint
    t0 = time.perf_counter()
    hosts = [W.host_case5(k) for k in range(nframe)]
    frameworks_s = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    items = []
    for host in hosts:
        prog = W.framework_program(host, host.notes['npoints'], images=True)
        items.append((prog, W.chempots_at(W.pressures_uneven(npress))))
    torch.cuda.synchronize()
    vext_s = time.perf_counter() - t0
    ub = UnevenBatch(items, m=5, damping=0.1, threshold=1e-8, nsteps=500)
    t0 = time.perf_counter()
    res = ub.solve()
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    grads = [BatchedAdjoint(p, tol=1e-10).gradients(r['rho']) for (p, _), r in zip(items, res)]
    torch.cuda.synchronize()
    adj_s = time.perf_counter() - t0
    nsolve = sum(len(r['N']) for r in res)
    npts = [int(np.prod(p.grid.npoints)) for p, _ in items]
    iters = np.concatenate([r['iters'] for r in res])
    return {'workload': '%d synthetic frameworks (grids %s..%s points) x %d pressures 0.1..50 bar, WBII + MFA-WDA, '
                        'periodic-image LJ' % (nframe, min(npts), max(npts), npress),
            'solves': nsolve, 'solve_s': solve_s, 'points_per_s': nsolve / solve_s,
            'all_converged': bool(all(np.all(r['converged']) for r in res)),
            'resolved_points': int(sum(len(r['resolved']) for r in res)),
            'iterations_mean': float(iters.mean()), 'iterations_max': int(iters.max()),
            'peak_concurrent_frameworks': ub.stats['peak_concurrent'], 'peak_bytes': ub.stats['peak_bytes'],
            'adjoint_s': adj_s, 'adjoint_vs_solve': adj_s / solve_s,
            'cg_iterations_max': int(max(g['iters'] for g in grads)), 'frameworks_s': frameworks_s,
            'vext_s': vext_s,
            'total_uptake_grad_norm': [float(np.linalg.norm(np.concatenate([g['dNtot_dsigma'], g['dNtot_depsilon']])))
                                       for g in grads][:4]}


def passes_only(shape, reps=5, nbatch=1):
    """Functional chain of one realistic density (ideal-gas start at 10 bar),
    on ``nbatch`` instances of it (one plan, every kernel launch covering
    all of them, as in the batched isotherm), each pass launched ``reps``
    times (the z pass each time after a seed and a forward evaluation, so
    that it reads weighted densities: the profile's k_zphi128 launches
    alternate between that evaluation's, which store nothing, and the pass's
    own), then a plain
    256 MB device-to-device copy (the reference for scaling FETCH_SIZE /
    WRITE_SIZE counters)."""
    import torch
    from cmmdft_amd import workloads as W
    from cmmdft_amd.eos import ModifiedBenedictWebbRubinEOS
    host, prog = build_program(shape)
    prog.fener.add_external_potential(cutoff=W.CUTOFF)
    from cmmdft_amd import units as U
    mu = ModifiedBenedictWebbRubinEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON).calculate_mu(W.TEMPERATURE,
                                                                                                  10 * U.bar)
    from cmmdft_amd.solver import _Engine
    eng = _Engine(prog.grid, prog.fener)
    rho = prog._set_initial_density(chempot=mu, temperature=W.TEMPERATURE).reshape(-1).contiguous()
    plan = eng.plan
    if nbatch > 1:
        from cmmdft_amd._lib import Plan
        from cmmdft_amd.functionals import configure_plan
        plan = Plan(prog.grid.npoints, prog.grid.cell.rvecs, nbatch=nbatch)
        configure_plan(plan, eng.hs, eng.wda, eng.mfa, eng.kT)
        rho = rho.repeat(nbatch).contiguous()
    plan.seed(rho, eng.V, False, eng.lam3)
    plan.derive()
    torch.cuda.synchronize()
    names = pass_names(chain_fields(plan))
    prep = zpass_input(plan, rho, eng.V, eng.lam3)
    for which in PASSES:
        for _ in range(reps):
            if which == 0:
                prep()   # the z pass's input: the weighted densities (see time_passes)
            plan.launch_pass(which)
    torch.cuda.synchronize()
    a = torch.empty(256 * 2 ** 20 // 8, dtype=torch.float64, device='cuda').fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    t = time_passes(plan, prep)
    print(json.dumps({'passes_only': True, 'grid': list(shape), 'nbatch': nbatch,
                      'pass_ms': {names[k]: v for k, v in t.items()},
                      'pass_bytes': {names[k]: nbatch * v
                                     for k, v in pass_bytes(*shape, chain_fields(plan)).items()},
                      'copy_bytes': a.numel() * 8}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--points', type=int, default=32)
    ap.add_argument('--mode', choices=['both', 'batched', 'plain'], default='both')
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--grid', type=int, default=NSIDE)
    ap.add_argument('--cpu-sample', type=int, default=4, help='oracle iterations timed (0 skips the CPU leg)')
    ap.add_argument('--no-adjoint', action='store_true', help='skip the forward-vs-adjoint timing')
    ap.add_argument('--no-uneven', action='store_true', help='skip the uneven-batch (16 frameworks) leg')
    ap.add_argument('--no-effpot', action='store_true', help='skip the non-spherical (CO2) effective-potential leg')
    ap.add_argument('--effpot-only', action='store_true', help='only the CO2 effective-potential leg (profiling)')
    ap.add_argument('--batch', type=int, default=1, help='instances per launch in --passes-only')
    ap.add_argument('--passes-only', action='store_true',
                    help='short run for hardware counters: the functional passes at the bench grid, then a 256 MB '
                         'device copy as the byte-count reference')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build runs on a single MI355X; --gpus %d is not supported'
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from cmmdft_amd import units as U, workloads as W
    from cmmdft_amd.solver import Anderson
    shape = (args.grid,) * 3
    if args.passes_only:
        return passes_only(shape, nbatch=args.batch)
    if args.effpot_only:
        print(json.dumps({'effective_potential': effpot_leg(shape)}))
        return
    mus = chempots(32)
    host, prog = build_program(shape)
    torch.cuda.synchronize()
    tv = time.perf_counter()
    prog.fener.add_external_potential(cutoff=W.CUTOFF)
    torch.cuda.synchronize()
    vext_ms = (time.perf_counter() - tv) * 1e3
    # the same once the device code is loaded (grid sum and clamp), on a
    # part that is then dropped again
    tv = time.perf_counter()
    prog.fener.add_external_potential(cutoff=W.CUTOFF)
    torch.cuda.synchronize()
    vext_warm_ms = (time.perf_counter() - tv) * 1e3
    prog.fener.parts.pop()
    prog.fener.part_names.pop()
    prog.fener.part_dict['ExtPot'] = prog.fener.parts[-1]
    solver = Anderson(prog, nsteps=500, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-8)
    prog.set_solver(solver)
    from cmmdft_amd.batch import BatchedAnderson
    mus = mus[:args.points]
    npt = len(mus)

    def plain_isotherm():
        its, cv = [], []
        for mu in mus:
            rho0 = prog._set_initial_density(chempot=mu, temperature=W.TEMPERATURE)
            N, rho, conv = solver.solve(mu, rho0)
            its.append(solver.curr_step)
            cv.append(conv)
        return its, cv

    batch = BatchedAnderson(prog, mus, m=5, damping=0.1, threshold=1e-8, nsteps=500)

    def batched_isotherm():
        res = batch.solve()
        return list(res['iters']), list(res['converged']), res

    plain_pps = batched_pps = None
    iters, convs = [], []
    if args.mode in ('both', 'plain'):
        for _ in range(args.warmup):
            plain_isotherm()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            its, cv = plain_isotherm()
            iters += its
            convs += cv
        torch.cuda.synchronize()
        dt_plain = time.perf_counter() - t0
        plain_pps = npt * args.steps / dt_plain
    biters, bconvs, resolved = [], [], 0
    if args.mode in ('both', 'batched'):
        for _ in range(args.warmup):
            batched_isotherm()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            its, cv, res = batched_isotherm()
            biters += its
            bconvs += cv
            resolved += len(res['resolved'])
        torch.cuda.synchronize()
        dt_batch = time.perf_counter() - t0
        batched_pps = npt * args.steps / dt_batch
    value = batched_pps if batched_pps is not None else plain_pps
    dt = (dt_batch if batched_pps is not None else dt_plain)
    if not iters:
        iters, convs = biters, bconvs
    points_per_s = value
    iters_per_point = float(np.mean(iters))

    # dominant pass: per-launch time x launches per iteration (2 paired y passes),
    # timed on the single-instance plan (solve one point first if only the batch ran)
    if plain_pps is None:
        solver.solve(mus[0], prog._set_initial_density(chempot=mus[0], temperature=W.TEMPERATURE))
    from cmmdft_amd.solver import _Engine
    eng = _Engine(prog.grid, prog.fener)
    rho_in = prog._set_initial_density(chempot=mus[0], temperature=W.TEMPERATURE).reshape(-1).contiguous()
    plan = prog.grid.plan()
    tp = time_passes(plan, zpass_input(plan, rho_in, eng.V, eng.lam3))
    per_iter_weight = {0: 1, 1: 1, 2: 1, 3: 2}
    nyf = chain_fields(plan)
    names = pass_names(nyf)
    pb = pass_bytes(*shape, nyf)
    # the same passes as the batched isotherm launches them: one launch per
    # pass covering all nb instances of the batch plan (every instance
    # active); the roofline is taken from these when the batch ran
    nbr, tpr = 1, tp
    if batched_pps is not None:
        bplan = batch._groups[0]['plan']
        bplan.set_active(None)
        nbr = bplan.nbatch
        tpr = time_passes(bplan, zpass_input(bplan, rho_in.repeat(nbr).contiguous(), eng.V, eng.lam3))
    dom = max(tpr, key=lambda w: tpr[w] * per_iter_weight[w])
    achieved = nbr * pb[dom] / (tpr[dom] * 1e-3) / 1e9
    npts = int(np.prod(shape))
    # real-field 3-D transforms per iteration: rho, the six weighted
    # densities, their six derivatives, c1 (12 without rho_bar's pair when
    # the WDA is applied in k space), each read and written once
    ntr = 12 if nyf == 4 else 14
    b_iter = npts * (16 * ntr + 24) + npts * 8 * (3 * 5 + 2)
    iter_gbps = b_iter * int(sum(int(v) for v in (biters if biters else iters))) / dt / 1e9

    adjoint = None if args.no_adjoint else adjoint_leg(shape)
    uneven = None if args.no_uneven else uneven_leg()
    effpot = None if args.no_effpot else effpot_leg(shape)

    cpu = None
    if args.cpu_sample > 0:
        V = prog.fener.part_dict['ExtPot'].potential
        cpu = cpu_baseline(V, shape, mus[len(mus) // 2], iters_per_point, args.cpu_sample)
        # the same on the host's cores (at most 16: one GPU's share of the box)
        ncores = min(16, len(os.sched_getaffinity(0)))
        if ncores > 1:
            cpu['multi_core'] = cpu_baseline(V, shape, mus[len(mus) // 2], iters_per_point, args.cpu_sample,
                                             cores=ncores)

    line = {
        'metric': METRIC, 'value': points_per_s, 'unit': 'points/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': dt / args.steps * 1e3, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': 'isotherm at %dx%dx%d: case-2 host (a=25.6 A, 276 LJ atoms), WBII FMT + MFA-WDA + '
                               'LJ ext. potential, Anderson m=5 damping 0.1 RIUE<1e-8, %d pressures 1e-2..1e2 bar '
                               'solved per step' % (shape + (npt,)),
                   'grid': list(shape), 'solver': 'anderson', 'points_per_step': npt,
                   'mode': 'batched' if batched_pps is not None else 'plain'},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': PEAK_HBM_GBPS, 'unit': 'GB/s',
                     'frac': achieved / PEAK_HBM_GBPS,
                     'traffic': (measured_traffic(names[dom], shape[0], nbr) if shape[0] == shape[1] == shape[2]
                                 else None),
                     'kernel': names[dom], 'instances_per_launch': nbr,
                     'avg_launch_ms': tpr[dom], 'algorithmic_bytes_per_launch': nbr * pb[dom]},
        'cpu_baseline': cpu,
        'plain_points_per_s': plain_pps,
        'batched_points_per_s': batched_pps,
        'batched_vs_plain': (batched_pps / plain_pps) if (batched_pps and plain_pps) else None,
        'batched_resolved_points': resolved,
        'batch_chunk': batch.chunk,
        'iterations_per_point': iters_per_point,
        'max_iterations_per_point': int(max(iters)),
        'iteration_algorithmic_GBps': iter_gbps,
        'all_converged': bool(all(convs) and all(bconvs)),
        'vext_ms': vext_ms,
        'vext_warm_ms': vext_warm_ms,
        'pass_ms': {names[k]: v for k, v in tp.items()},
        'pass_ms_batched': ({names[k]: v for k, v in tpr.items()} if nbr > 1 else None),
        'pass_GBps_batched': ({names[k]: nbr * pb[k] / (v * 1e-3) / 1e9 for k, v in tpr.items()}
                              if nbr > 1 else None),
        'adjoint': adjoint,
        'uneven_batch': uneven,
        'effective_potential': effpot,
    }
    print(json.dumps(line))


if __name__ == '__main__':
    main()
