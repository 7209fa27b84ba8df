"""Mean-field (MFA) functional on the device against the oracle's
restatement of MFAFunctional (cmmdft/functionals.py:770-939): the pair
potential, its kernel table (generated with the Lanczos factor, loaded
without), with and without supercell tail corrections, derive/value, and
converged solutions with the reference notebook's functional set
(atWBII + MFA + correlation WDA + external potential)."""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, rho_rel_err

pytestmark = pytest.mark.gpu

RC = 12 * U.angstrom


def _grid(shape, rvecs):
    from cmmdft_amd.system import Grid
    return Grid(rvecs, shape)


def _density(shape, seed=4):
    rng = np.random.default_rng(seed)
    return 2e-3 / U.angstrom ** 3 * (1 + 0.5 * rng.random(shape))


CASES = [((24, 24, 24), 24.0, False), ((20, 18, 25), 26.0, False), ((16, 15, 12), 11.0, True),
         ((12, 10, 9), 9.5, True)]


@pytest.mark.parametrize('shape,edge,tail', CASES)
def test_generated_potential_derive_and_value(shape, edge, tail):
    import torch
    from cmmdft_amd.functionals import MFAFunctional
    rv = np.diag([edge, edge * 1.05, edge * 0.97]) * U.angstrom
    og = O.Grid(rv, shape)
    om = O.MFA(og, tailcorrections=tail)
    om.generate_potential_lj(W.GUEST_SIGMA, W.GUEST_EPSILON, cutoff=RC)
    m = MFAFunctional(_grid(shape, rv), tailcorrections=tail)
    m.generate_potential_lj(W.GUEST_SIGMA, W.GUEST_EPSILON, cutoff=RC)
    # the (super)grid potential
    assert m.potential.shape == om.potential.shape
    assert np.abs(m.potential - om.potential).max() <= 1e-13 * np.abs(om.potential).max()
    assert abs(m.compute_vdw_a() - om.compute_vdw_a()) <= 1e-12 * abs(om.compute_vdw_a())
    rho = _density(shape)
    kr = og.fft(rho)
    c_ref = om.derive(kr)
    c = m.derive(torch.from_numpy(rho).cuda()).cpu().numpy()
    assert np.abs(c - c_ref).max() <= 1e-12 * np.abs(c_ref).max()
    # reference-convention spectrum in, numpy out
    assert np.abs(m.derive(kr) - c_ref).max() <= 1e-12 * np.abs(c_ref).max()
    F_ref = om.value(kr)
    assert abs(m.value(torch.from_numpy(rho).cuda()) - F_ref) <= 1e-12 * abs(F_ref)
    loc = m.value(torch.from_numpy(rho).cuda(), local=True).cpu().numpy()
    assert np.abs(loc - om.value(kr, local=True)).max() <= 1e-12 * np.abs(om.value(kr, local=True)).max()


@pytest.mark.parametrize('tail', [False, True])
def test_loaded_potential_has_no_lanczos(tail, tmp_path):
    """load_potential: the kernel is fft(U) without the Lanczos factor."""
    import torch
    from cmmdft_amd.functionals import MFAFunctional
    shape = (16, 16, 16) if not tail else (10, 12, 8)
    rv = np.eye(3) * (20.0 if not tail else 9.0) * U.angstrom
    og = O.Grid(rv, shape)
    om = O.MFA(og, tailcorrections=tail)
    om.generate_potential_lj(W.GUEST_SIGMA, W.GUEST_EPSILON, cutoff=RC)
    rng = np.random.default_rng(2)
    pot = om.potential * (1 + 0.1 * rng.random(om.potential.shape))
    fn = tmp_path / 'mfa.npy'
    np.save(fn, pot)
    om.load_potential(pot)
    m = MFAFunctional(_grid(shape, rv), tailcorrections=tail)
    m.load_potential(fn)
    rho = _density(shape, 7)
    c_ref = om.derive(og.fft(rho))
    c = m.derive(torch.from_numpy(rho).cuda()).cpu().numpy()
    assert np.abs(c - c_ref).max() <= 1e-12 * np.abs(c_ref).max()
    # dump writes the (super)grid potential back
    m.dump_potential(str(tmp_path / 'sub' / 'out.npy'))
    assert np.array_equal(np.load(tmp_path / 'sub' / 'out.npy'), pot)


def _notebook_program(host, shape, tail):
    from cmmdft_amd.program import Program
    from cmmdft_amd.system import NanoporousHost, SphericalLJGuest
    prog = Program(write_files=False)
    prog.set_system(NanoporousHost('host', host, host.ff),
                    SphericalLJGuest('CH4', W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    prog.set_grid(npoints=list(shape))
    prog.init_free_energy(T)
    prog.fener.add_hard_sphere('atWBII')
    prog.fener.add_mean_field(tailcorrections=tail)
    prog.fener.add_correlation_wda_lj()
    prog.fener.add_external_potential(cutoff=W.CUTOFF)
    return prog


def _notebook_oracle(host, shape, tail):
    og = O.Grid(host.rvecs, shape)
    R = O.bh_radius(W.GUEST_SIGMA, W.GUEST_EPSILON, T)
    lam = O.thermal_wavelength(W.GUEST_MASS, T)
    sig, eps = host.atom_params()
    V = O.external_potential(og.points[..., :3], sig, eps, host.pos, host.rvecs, W.GUEST_SIGMA, W.GUEST_EPSILON)
    mfa = O.MFA(og, tailcorrections=tail)
    mfa.generate_potential_lj(W.GUEST_SIGMA, W.GUEST_EPSILON, cutoff=RC)
    m, s, e = W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON
    corr = O.SumOfEOS(m, [O.MBWREOS(m, s, e), O.CarnahanStarlingEOS(m, R), O.MFAEOS(m, s, e, a=mfa.compute_vdw_a())],
                      [1, -1, -1])
    parts = [O.HardSphere(og, R, 'atWBII', T), mfa, O.WDAV(og, R, corr, T), O.ExtPot(og, V)]
    return og, parts, 1 / (U.boltzmann * T), lam, parts[-1].potential


@pytest.mark.parametrize('tail', [False, True])
def test_notebook_functional_solution_matches_oracle(tail):
    """atWBII + MFA (+ tail corrections) + correlation WDA + LJ external
    potential: GPU Anderson against the oracle's static Picard (rho, N and
    Omega within 1e-8)."""
    from cmmdft_amd.solver import Anderson, grand_potential
    host = W.host_case1()
    shape = (20, 20, 20)
    og, parts, beta, lam, V = _notebook_oracle(host, shape, tail)
    mu = mu_at(10.0)
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    s = O.Picard(og, parts, beta, lam, nsteps=5000, alpha_mix=0.05, method='static', criterion='RIUE',
                 threshold=1e-13)
    Nref, rho_ref, conv = s.solve(mu, rho0)
    assert conv
    omega_ref = s.omega(rho_ref, og.fft(rho_ref))
    prog = _notebook_program(host, shape, tail)
    assert prog.fener.part_names == ['HardSphere', 'MFA', 'WDA-V', 'ExtPot']
    prog.set_solver(Anderson(prog, nsteps=1000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                             threshold=1e-13))
    N, rho, conv = prog.solve(mu)
    assert conv
    rel, mism = rho_rel_err(rho, rho_ref)
    assert rel <= 1e-8, rel
    assert abs(N - Nref) <= 1e-8 * Nref
    om = grand_potential(prog.fener, mu, rho)
    assert abs(om - omega_ref) <= 1e-8 * abs(omega_ref), (om, omega_ref)


def test_notebook_functional_batch_equals_plain():
    import torch
    from cmmdft_amd.batch import BatchedAnderson
    from cmmdft_amd.solver import Anderson
    host = W.host_case1()
    prog = _notebook_program(host, (20, 20, 20), True)
    eos = O.MBWREOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON)
    mus = np.array([eos.calculate_mu(T, p) for p in W.pressures_isotherm(4)])
    b = BatchedAnderson(prog, mus, m=5, damping=0.1, threshold=1e-10, nsteps=400)
    res = b.solve()
    assert res['converged'].all(), res['status']
    for i, mu in enumerate(mus):
        if i in res['resolved']:
            continue
        s = Anderson(prog, nsteps=400, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-10)
        r0 = prog._set_initial_density(chempot=mu, temperature=T)
        N, rho, conv = s.solve(mu, r0)
        assert conv and s.curr_step == res['iters'][i]
        assert torch.equal(rho.reshape(-1), res['rho'][i].reshape(-1))
    b.close()


def test_adjoint_rejects_mfa():
    from cmmdft_amd.adjoint import AdjointSolver
    prog = _notebook_program(W.host_case1(), (16, 16, 16), False)
    with pytest.raises(NotImplementedError):
        AdjointSolver(prog)
