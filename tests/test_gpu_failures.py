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
