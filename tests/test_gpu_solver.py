"""Converged Euler-Lagrange solutions of the HIP solvers against the
oracle's restatement of the reference solvers (cmmdft/solver.py).  Fixed
points are solver-independent, so the GPU Anderson (exact simplex LSQ) is also
checked against the oracle's Picard iteration."""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, oracle_system, product_program, rho_rel_err

pytestmark = pytest.mark.gpu

SHAPE = (24, 24, 24)


@pytest.fixture(scope='module')
def case1_ref():
    host = W.host_case1()
    og, parts, beta, lam, V = oracle_system(host, SHAPE, 'WBII')
    mu = mu_at(10.0)
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    s = O.Picard(og, parts, beta, lam, nsteps=5000, alpha_mix=0.05, method='static', criterion='RIUE',
                 threshold=1e-13)
    N, rho, conv = s.solve(mu, rho0)
    assert conv
    omega = s.omega(rho, og.fft(rho))
    return host, mu, N, rho, omega


def _check(prog, mu, N, rho, Nref, rho_ref, omega_ref, same_family=False):
    from cmmdft_amd.solver import grand_potential
    rel, mism = rho_rel_err(rho, rho_ref)
    assert rel <= 1e-8, rel
    if same_family:
        assert mism == 0
    assert abs(N - Nref) <= 1e-8 * abs(Nref)
    om = grand_potential(prog.fener, mu, rho)
    assert abs(om - omega_ref) <= 1e-8 * abs(omega_ref), (om, omega_ref)


def test_picard_static_matches_oracle(case1_ref):
    from cmmdft_amd.solver import Picard
    host, mu, Nref, rho_ref, omega_ref = case1_ref
    prog = product_program(host, SHAPE, 'WBII')
    prog.set_solver(Picard(prog, nsteps=5000, alpha_mix=0.05, method='static', criterion='RIUE', threshold=1e-13))
    N, rho, conv = prog.solve(mu)
    assert conv
    _check(prog, mu, N, rho, Nref, rho_ref, omega_ref, same_family=True)


@pytest.mark.parametrize('method', ['anderson', 'hybridanderson'])
def test_anderson_matches_oracle(case1_ref, method):
    from cmmdft_amd.solver import Anderson
    host, mu, Nref, rho_ref, omega_ref = case1_ref
    prog = product_program(host, SHAPE, 'WBII')
    prog.set_solver(Anderson(prog, nsteps=1000, method=method, m=5, damping=0.1, criterion='RIUE', threshold=1e-13))
    N, rho, conv = prog.solve(mu)
    assert conv, prog.solver.curr_step
    _check(prog, mu, N, rho, Nref, rho_ref, omega_ref)


def test_oracle_anderson_same_fixed_point(case1_ref):
    """The oracle's own SLSQP Anderson lands on the Picard fixed point too."""
    host, mu, Nref, rho_ref, omega_ref = case1_ref
    og, parts, beta, lam, V = oracle_system(host, SHAPE, 'WBII')
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    a = O.Anderson(og, parts, beta, lam, nsteps=1000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                   threshold=1e-13)
    N, rho, conv = a.solve(mu, rho0)
    assert conv
    rel, _ = rho_rel_err(rho, rho_ref)
    assert rel <= 1e-8


def test_hybrid_picard_matches_oracle(case1_ref):
    from cmmdft_amd.solver import Picard
    host, mu, Nref, rho_ref, omega_ref = case1_ref
    prog = product_program(host, SHAPE, 'WBII')
    prog.set_solver(Picard(prog, nsteps=3000, alpha_mix=0.1, method='hybrid', criterion='RIUE', threshold=1e-13))
    N, rho, conv = prog.solve(mu)
    assert conv, prog.solver.curr_step
    _check(prog, mu, N, rho, Nref, rho_ref, omega_ref)


def test_rosenfeld_anderson_vs_oracle_picard():
    from cmmdft_amd.solver import Anderson
    host = W.host_case1()
    shape = (20, 20, 20)
    og, parts, beta, lam, V = oracle_system(host, shape, 'FMT')
    mu = mu_at(10.0)
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    s = O.Picard(og, parts, beta, lam, nsteps=5000, alpha_mix=0.05, method='static', criterion='RIUE',
                 threshold=1e-13)
    Nref, rho_ref, conv = s.solve(mu, rho0)
    assert conv
    prog = product_program(host, shape, 'FMT')
    prog.set_solver(Anderson(prog, nsteps=1000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                             threshold=1e-13))
    N, rho, conv = prog.solve(mu)
    assert conv
    rel, mism = rho_rel_err(rho, rho_ref)
    assert rel <= 1e-8
    assert abs(N - Nref) <= 1e-8 * Nref


@pytest.mark.parametrize('version', ['atWBII', 'tFMT'])
def test_tensor_anderson_vs_oracle_picard(version):
    """The tensor variants (atWBII is the reference's default hard-sphere
    functional): GPU Anderson against the oracle's static Picard."""
    from cmmdft_amd.solver import Anderson, grand_potential
    host = W.host_case1()
    shape = (20, 20, 20)
    og, parts, beta, lam, V = oracle_system(host, shape, version)
    mu = mu_at(10.0)
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    s = O.Picard(og, parts, beta, lam, nsteps=5000, alpha_mix=0.05, method='static', criterion='RIUE',
                 threshold=1e-13)
    Nref, rho_ref, conv = s.solve(mu, rho0)
    assert conv
    omega_ref = s.omega(rho_ref, og.fft(rho_ref))
    prog = product_program(host, shape, version)
    prog.set_solver(Anderson(prog, nsteps=1000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                             threshold=1e-13))
    N, rho, conv = prog.solve(mu)
    assert conv
    rel, mism = rho_rel_err(rho, rho_ref)
    assert rel <= 1e-8, rel
    assert abs(N - Nref) <= 1e-8 * Nref
    om = grand_potential(prog.fener, mu, rho)
    assert abs(om - omega_ref) <= 1e-8 * abs(omega_ref), (om, omega_ref)


def test_consecutive_solves_do_not_leak_state():
    """Two solves at different chemical potentials on one program: the second
    must not reuse the first one's device state (fugacity, history, damping),
    so solving mu2 after mu1 equals solving mu2 on a fresh program."""
    from cmmdft_amd.solver import Anderson
    import torch
    host = W.host_case1()
    shape = (16, 16, 16)
    mu1, mu2 = mu_at(1.0), mu_at(30.0)
    prog = product_program(host, shape, 'WBII')
    s = Anderson(prog, nsteps=300, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-11)
    prog.set_solver(s)
    prog.solve(mu1)
    N2, rho2, c2 = prog.solve(mu2)
    it2 = s.curr_step
    fresh = product_program(host, shape, 'WBII')
    s2 = Anderson(fresh, nsteps=300, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-11)
    fresh.set_solver(s2)
    N2f, rho2f, c2f = fresh.solve(mu2)
    assert c2 and c2f and it2 == s2.curr_step
    assert np.array_equal(rho2, rho2f) and N2 == N2f
