"""Config 2 at its stated size, the workload the benchmark is quoted on:
the case-2 host (cubic a = 25.6 A, 276 LJ atoms of 4 types), 128^3 grid,
united-atom methane at 300 K and 10 bar, White-Bear II FMT + mean-field WDA
+ LJ external potential, Anderson m = 5, damping 0.1.

* The plain HIP solver against the oracle's Anderson solve of the same
  state point (both to RIUE < 1e-13): V_ext <= 1e-10, rho <= 1e-8 max
  relative over rho > 1e-10 A^-3 with equal zero masks, N and Omega <= 1e-8.
* The benchmark's isotherm (32 pressures 1e-2..1e2 bar) batched: every
  point converges, uptake rises with pressure, the low-pressure end is the
  Henry-law limit of the same potential, and three state points equal the
  plain solver's bit for bit.
"""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, oracle_system, product_program, rho_rel_err

pytestmark = pytest.mark.gpu

SHAPE = (128, 128, 128)


@pytest.fixture(scope='module')
def prog():
    return product_program(W.host_case2(), SHAPE, 'WBII')


def test_config2_plain_solver_matches_oracle(prog):
    from cmmdft_amd.solver import Anderson, grand_potential
    host = W.host_case2()
    mu = mu_at(10.0)
    og, parts, beta, lam, V = oracle_system(host, SHAPE, 'WBII')
    # external potential (unclamped points)
    Vp = prog.fener.part_dict['ExtPot'].potential
    free = V < 1e4 * U.kjmol
    scale = np.abs(V[free]).max()
    assert np.abs(Vp - V)[free].max() <= 1e-10 * scale
    big = free & (np.abs(V) > 1e-6 * scale)
    assert (np.abs(Vp - V)[big] / np.abs(V[big])).max() <= 1e-10
    assert np.array_equal(Vp >= 1e4 * U.kjmol, ~free)
    # the oracle's solve (scipy FFTs, SLSQP mixing)
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    s = O.Anderson(og, parts, beta, lam, nsteps=400, method='anderson', m=5, damping=0.1, criterion='RIUE',
                   threshold=1e-13)
    Nref, rho_ref, conv = s.solve(mu, rho0)
    assert conv
    omega_ref = s.omega(rho_ref, og.fft(rho_ref))
    solver = Anderson(prog, nsteps=400, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-13)
    N, rho, conv = solver.solve(mu, prog._set_initial_density(chempot=mu, temperature=T))
    assert conv
    rho = rho.cpu().numpy()
    rel, mism = rho_rel_err(rho, rho_ref)
    assert rel <= 1e-8 and mism == 0, (rel, mism)
    assert abs(N - Nref) <= 1e-8 * Nref
    om = grand_potential(prog.fener, mu, rho)
    assert abs(om - omega_ref) <= 1e-8 * abs(omega_ref)


def test_config2_isotherm_batched_properties(prog):
    import torch
    from cmmdft_amd.batch import BatchedAnderson
    from cmmdft_amd.solver import Anderson
    mus = W.chempots_at(W.pressures_isotherm(32))
    b = BatchedAnderson(prog, mus, m=5, damping=0.1, threshold=1e-8, nsteps=500)
    res = b.solve()
    assert np.all(res['converged'])
    assert np.all(np.diff(res['N']) > 0)
    # Henry limit at 1e-2 bar: the ideal-gas Boltzmann integral of the same potential
    V = prog.fener.part_dict['ExtPot'].potential
    beta, lam3 = prog.fener.beta, prog.fener.wavelength ** 3
    N_ideal = float(np.sum(np.exp(beta * (mus[0] - V)))) / lam3 * prog.grid.dr
    assert abs(res['N'][0] / N_ideal - 1) < 5e-3
    solver = Anderson(prog, nsteps=500, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-8)
    for i in (0, 16, 31):
        if i in res['resolved']:
            continue
        N, rho, conv = solver.solve(mus[i], prog._set_initial_density(chempot=mus[i], temperature=T))
        assert conv and N == res['N'][i] and solver.curr_step == res['iters'][i]
        assert torch.equal(rho.reshape(-1), res['rho'][i].reshape(-1))
    b.close()
