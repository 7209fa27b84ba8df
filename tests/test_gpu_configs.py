"""The measured configurations at their stated sizes.

Config 1 (the reference's own CPU case): case-1 host, 48^3, Picard static
with mixing 0.05 to max|drho| < 1e-8 A^-3, WBII + mean-field WDA + LJ: the
HIP solver against the oracle running the same iteration.
Config 3 (batched isotherm): case-1 host at 96^3, the 32 pressures of the
isotherm log-spaced 1e-2..1e2 bar: batched = plain bit for bit on a subset,
N(p) increasing, and the low-pressure end against the ideal-gas Boltzmann
integral (the Henry-law limit)."""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, oracle_system, product_program, rho_rel_err

pytestmark = pytest.mark.gpu


def test_config1_picard_maxdrho_matches_oracle():
    from cmmdft_amd.solver import Picard
    host = W.host_case1()
    shape = (48, 48, 48)
    mu = mu_at(10.0)
    og, parts, beta, lam, V = oracle_system(host, shape, 'WBII')
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    s = O.Picard(og, parts, beta, lam, nsteps=20000, alpha_mix=0.05, method='static', criterion='maxdrho',
                 threshold=1e-8)
    Nref, rho_ref, conv = s.solve(mu, rho0)
    assert conv
    prog = product_program(host, shape, 'WBII')
    solver = Picard(prog, nsteps=20000, alpha_mix=0.05, method='static', criterion='maxdrho', threshold=1e-8)
    prog.set_solver(solver)
    N, rho, conv = prog.solve(mu)
    assert conv
    # the same contraction run for the same number of steps from the same start
    assert solver.curr_step == s.curr_step
    rel, mism = rho_rel_err(rho, rho_ref)
    assert rel <= 1e-8 and mism == 0
    assert abs(N - Nref) <= 1e-8 * Nref


@pytest.fixture(scope='module')
def config3():
    from cmmdft_amd.batch import BatchedAnderson
    host = W.host_case1()
    prog = product_program(host, (96, 96, 96), 'WBII')
    mus = W.chempots_at(W.pressures_isotherm(32))
    res = BatchedAnderson(prog, mus, m=5, damping=0.1, threshold=1e-8, nsteps=500).solve()
    return prog, mus, res


def test_config3_batched_isotherm(config3):
    prog, mus, res = config3
    assert np.all(res['converged']) and not res['resolved']
    assert np.all(np.diff(res['N']) > 0)
    # Henry limit at 1e-2 bar: ideal-gas Boltzmann integral of the same
    # potential, corrected to first order by the bulk excess chemical potential
    V = prog.fener.part_dict['ExtPot'].potential
    beta, lam3 = prog.fener.beta, prog.fener.wavelength ** 3
    N_ideal = float(np.sum(np.exp(beta * (mus[0] - V)))) / lam3 * prog.grid.dr
    assert abs(res['N'][0] / N_ideal - 1) < 5e-3


def test_config3_batched_equals_plain(config3):
    from cmmdft_amd.solver import Anderson
    prog, mus, res = config3
    solver = Anderson(prog, nsteps=500, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-8)
    prog.set_solver(solver)
    for i in (0, 13, 31):
        rho0 = prog._set_initial_density(chempot=mus[i], temperature=T)
        N, rho, conv = solver.solve(mus[i], rho0)
        assert conv and N == res['N'][i] and solver.curr_step == res['iters'][i]
