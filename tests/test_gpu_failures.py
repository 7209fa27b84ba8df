"""Failure handling of the solvers (solver.py:283-345, program.py:315-329),
forced from host data: every kernel runs to completion on the (non-finite)
values it is given, and the host control flow reacts to the device's
non-finite counts as the reference does to np.isfinite.

* a non-finite initial density or a Boltzmann factor that overflows:
  FloatingPointError in every phase of the correction-factor retry loop,
  then NoSolutionError;
* the step limit: converged = False with the last density;
* an infinite mixing parameter (static Picard): the same retry path;
* a batched instance with a non-finite start: status 4, its plain re-solve
  fails too, and the other instances' results are unchanged bit for bit;
* a solver cascade in which every solver fails: None and a warning.
"""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, product_program

pytestmark = pytest.mark.gpu

SHAPE = (16, 16, 16)


@pytest.fixture(scope='module')
def prog():
    return product_program(W.host_case1(), SHAPE, 'WBII')


def _nan_start(prog, mu):
    r0 = prog._set_initial_density(chempot=mu, temperature=T).cpu().numpy().copy()
    r0[3, 4, 5] = np.nan
    return r0


@pytest.mark.parametrize('method', ['anderson', 'hybridanderson'])
def test_nonfinite_start_raises_no_solution(prog, method):
    from cmmdft_amd.solver import Anderson, NoSolutionError
    mu = mu_at(10.0)
    s = Anderson(prog, nsteps=50, method=method, m=5, damping=0.1, criterion='RIUE', threshold=1e-8)
    with pytest.raises(NoSolutionError):
        s.solve(mu, _nan_start(prog, mu))
    # three phases: correction factors 1, 1/2, 1/4, each raising FloatingPointError
    assert s.iphase == 3
    assert s.correction_factor == 1


def test_overflowing_boltzmann_factor_raises_no_solution():
    """A deep well at one grid point (-1e4 hartree): exp(-beta C1)
    overflows there from the first update."""
    from cmmdft_amd.solver import Picard, NoSolutionError
    host = W.host_case1()
    p = product_program(host, SHAPE, 'WBII')
    ep = p.fener.part_dict['ExtPot']
    V = ep.potential.copy()
    V[2, 2, 2] = -1e4
    p.fener.remove_part('ExtPot')
    p.fener.add_external_potential(potential=V)
    mu = mu_at(10.0)
    r0 = p._set_initial_density(chempot=mu, temperature=T)
    r0[2, 2, 2] = 1e-3      # the ideal-gas start itself would already be infinite there
    s = Picard(p, nsteps=20, method='hybrid', criterion='RIUE', threshold=1e-8)
    with pytest.raises(NoSolutionError):
        s.solve(mu, r0)
    assert s.iphase == 3


def test_infinite_mixing_parameter_raises_no_solution(prog):
    from cmmdft_amd.solver import Picard, NoSolutionError
    mu = mu_at(10.0)
    s = Picard(prog, nsteps=20, alpha_mix=float('inf'), method='static', criterion='RIUE', threshold=1e-8)
    with pytest.raises(NoSolutionError):
        s.solve(mu, prog._set_initial_density(chempot=mu, temperature=T))


@pytest.mark.parametrize('nsteps', [1, 3])
def test_step_limit_returns_unconverged(prog, nsteps):
    from cmmdft_amd.solver import Anderson
    mu = mu_at(10.0)
    s = Anderson(prog, nsteps=nsteps, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-12)
    N, rho, conv = s.solve(mu, prog._set_initial_density(chempot=mu, temperature=T))
    assert conv is False and s.curr_step == nsteps
    assert np.isfinite(N) and bool(rho.isfinite().all())
    assert abs(N - float(rho.sum()) * prog.grid.dr) <= 1e-10 * N


def test_batched_nonfinite_instance_is_isolated(prog):
    import torch
    from cmmdft_amd.batch import BatchedAnderson
    mus = W.chempots_at(W.pressures_isotherm(4))
    r0 = torch.stack([prog._set_initial_density(chempot=m, temperature=T) for m in mus])
    good = BatchedAnderson(prog, mus, m=5, damping=0.1, threshold=1e-10, nsteps=300).solve(r0.clone())
    bad0 = r0.clone()
    bad0[1, 3, 4, 5] = float('nan')
    bad = BatchedAnderson(prog, mus, m=5, damping=0.1, threshold=1e-10, nsteps=300).solve(bad0)
    assert bad['status'][1] == 4 and not bad['converged'][1] and 1 in bad['resolved']
    assert np.isnan(bad['N'][1])
    for i in (0, 2, 3):
        assert bad['status'][i] == 1 and bad['N'][i] == good['N'][i]
        assert bad['iters'][i] == good['iters'][i]
        assert torch.equal(bad['rho'][i], good['rho'][i])


def test_cascade_where_every_solver_fails(prog):
    from cmmdft_amd.solver import Anderson, Picard
    mu = mu_at(10.0)
    solvers = [Anderson(prog, nsteps=20, method='hybridanderson', criterion='RIUE', threshold=1e-8),
               Picard(prog, nsteps=20, method='hybrid', criterion='RIUE', threshold=1e-8)]
    with pytest.warns(RuntimeWarning, match='All solvers failed'):
        out = prog.cascade_solver(solvers, mu, Ninit=_nan_start(prog, mu))
    assert out is None
    # the cascade falls through to the next solver when one merely stops at its step limit
    solvers = [Anderson(prog, nsteps=2, method='anderson', damping=0.1, criterion='RIUE', threshold=1e-12),
               Anderson(prog, nsteps=400, method='anderson', damping=0.1, criterion='RIUE', threshold=1e-9)]
    N, rho, conv = prog.cascade_solver(solvers, mu)
    assert conv and prog.solver is solvers[1]


# --- an Anderson density whose Boltzmann factor overflows ----------------
# (solver.py:561-574: update_rho takes the hybrid step from the step's
# starting density instead).  The overflow is planted from the host on both
# sides: the GPU solver's _force_nonfinite and the oracle's force_nonfinite
# treat the Anderson density of the given steps as having a non-finite
# Boltzmann factor; every kernel runs on finite values.

@pytest.mark.parametrize('method,m,j', [('anderson', 5, 5), ('anderson', 1, 4), ('hybridanderson', 5, 2)])
def test_nonfinite_anderson_density_takes_the_hybrid_step(method, m, j):
    """The Anderson density of step K (the j-th Anderson step) has a
    non-finite Boltzmann factor: the GPU solver notices it at the next
    residual pass and takes step K again as a hybrid step (history, ring
    slots, damping and tracking back at step K).  Step K's density is then
    the oracle's hybrid step (update_rho_hybrid, solver.py:403-457) from the
    same starting density, the steps before K are untouched, and the solve
    reaches the fixed point of the unforced one (and, m = 5, of the oracle's
    solve with the same planted overflow).  The GPU's exact simplex
    coefficients and the oracle's SLSQP ones give different Anderson
    trajectories, so whole trajectories are not compared; with m = 1 the
    reference's adaptive damping reads a second history row that does not
    exist (solver.py:537-544), so that case has no oracle solve."""
    from cmmdft_amd.solver import Anderson
    from helpers import oracle_system, rho_rel_err
    host = W.host_case1()
    og, parts, beta, lam, V = oracle_system(host, SHAPE, 'WBII')
    mu = mu_at(10.0)
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    prog = product_program(host, SHAPE, 'WBII')
    # m = 1 is damped Picard mixing: more steps
    nst = 400 if m > 1 else 4000
    kw = dict(method=method, m=m, damping=0.1, criterion='RIUE', threshold=1e-10, track_history=True)
    s0 = Anderson(prog, nsteps=nst, **kw)
    N0, r0, c0 = s0.solve(mu, rho0)
    K = s0.anderson_log[j][0]
    s = Anderson(prog, nsteps=nst, **kw)
    s._force_nonfinite = (K,)
    N, rho, conv = s.solve(mu, rho0)
    assert c0 and conv, (c0, conv, s0.curr_step, s.curr_step)
    assert s.redone_steps == [K] and s0.redone_steps == []
    assert np.array_equal(s.history[1:K, 0], s0.history[1:K, 0])
    assert abs(s.history[K, 0] - s0.history[K, 0]) > 1e-6 * abs(s0.history[K, 0])
    # step K against the oracle's hybrid step from the GPU's density before it
    sK = Anderson(prog, nsteps=K - 1, **kw)
    _, rK, _ = sK.solve(mu, rho0)
    op = O.Picard(og, parts, beta, lam, nsteps=1, alpha_mix=0.1, method='hybrid', criterion='RIUE',
                  threshold=1e-10)
    op._initiate(mu)
    kK = og.fft(rK)
    op.omega0 = op.omega(rK, kK)
    rH, _, _ = op.update_rho_hybrid(rK, kK, op.c1(rK, kK))
    NH = og.integrate(rH).real
    assert abs(s.history[K, 0] - NH) <= 1e-8 * abs(NH), (s.history[K, 0], NH)
    rel, mism = rho_rel_err(rho, r0)
    assert rel <= 1e-8 and mism == 0, rel
    if m > 1:
        o = O.Anderson(og, parts, beta, lam, nsteps=400, method=method, m=m, damping=0.1, criterion='RIUE',
                       threshold=1e-10)
        o.force_nonfinite = (K,)
        No, ro, co = o.solve(mu, rho0)
        assert co
        rel, mism = rho_rel_err(rho, ro)
        assert rel <= 1e-8 and mism == 0, rel
        assert abs(N - No) <= 1e-8 * abs(No)


def test_nonfinite_density_at_convergence_is_not_accepted():
    """The Anderson density that meets the criteria has a non-finite
    Boltzmann factor: it is not returned; that step is taken again as a
    hybrid step and the solve goes on (or stops, when the hybrid density
    meets the criteria too) at the same fixed point.  The tracking file
    keeps one line per step."""
    import os
    import tempfile
    from cmmdft_amd.solver import Anderson
    from helpers import rho_rel_err
    prog = product_program(W.host_case1(), SHAPE, 'WBII')
    mu = mu_at(10.0)
    r0 = prog._set_initial_density(chempot=mu, temperature=T)
    s0 = Anderson(prog, nsteps=400, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-10,
                  track_history=True)
    N0, rho0, c0 = s0.solve(mu, r0)
    Kc = s0.curr_step
    s0_hist_Kc = s0.history[Kc, 0]
    with tempfile.TemporaryDirectory() as d:
        prog.fener.init_tracking(os.path.join(d, 'conv.txt'), rewrite=True)
        s = Anderson(prog, nsteps=400, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-10,
                     track_history=True)
        s._force_nonfinite = (Kc,)
        N, rho, c = s.solve(mu, r0)
        lines = open(os.path.join(d, 'conv.txt'), encoding='utf-8').read().splitlines()
        prog.fener.fn_tracking = None
    # the density returned is the hybrid step's (which may meet the
    # criteria at the same step), not the Anderson one
    assert c0 and c and s.redone_steps == [Kc] and s.curr_step >= Kc
    assert s.history[Kc, 0] != s0_hist_Kc
    assert len(lines) == s.curr_step + 2
    assert [int(l.split()[1]) for l in lines[1:]] == list(range(s.curr_step + 1))
    rel, mism = rho_rel_err(rho.cpu().numpy(), rho0.cpu().numpy())
    assert rel <= 1e-8 and mism == 0, rel
