"""Real frameworks read from the reference's own input files (ZIF-8, cubic
with a slightly skewed cell; chabazite, hexagonal) through
NanoporousHost(name, chk, par): the device LJ external potential and
converged densities against the oracle."""
import os

import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, oracle_system, rho_rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'inputs')
FILES = {'ZIF8': ('zif8.chk', 'zif8_pars_vdw_uff.txt'), 'CHA': ('cha.chk', 'cha_pars.txt')}


def _host(name):
    from cmmdft_amd.system import NanoporousHost
    chk, par = FILES[name]
    return NanoporousHost(name, os.path.join(HERE, chk), os.path.join(HERE, par))


def _as_synthetic(h):
    """The same framework as the oracle helpers take it."""
    return W.SyntheticHost(np.asarray(h.cell.rvecs), h.mol.pos.copy(), np.asarray(h.mol.ffatype_ids),
                           list(h.mol.ffatypes), dict(h.ff), name=h.name)


@pytest.mark.parametrize('name,shape', [('ZIF8', (32, 32, 32)), ('ZIF8', (27, 30, 25)), ('CHA', (28, 28, 30))])
def test_real_framework_external_potential(name, shape):
    from cmmdft_amd.system import Grid
    from cmmdft_amd.extpot import get_external_potential
    h = _host(name)
    s = _as_synthetic(h)
    og = O.Grid(s.rvecs, shape)
    sig, eps = s.atom_params()
    Vr = O.external_potential(og.points[..., :3], sig, eps, s.pos, s.rvecs, W.GUEST_SIGMA, W.GUEST_EPSILON)
    V = get_external_potential(Grid(h.cell.rvecs, shape), h.ff, W.GUEST_SIGMA, W.GUEST_EPSILON, h.mol)
    scale = np.abs(Vr).max()
    big = np.abs(Vr) > 1e-6 * scale
    assert np.abs(V - Vr).max() <= 1e-10 * scale
    assert (np.abs(V - Vr)[big] / np.abs(Vr[big])).max() <= 1e-10


@pytest.mark.parametrize('name,shape,version,p_bar,tol_dilute', [('ZIF8', (24, 24, 24), 'WBII', 10.0, 1e-8),
                                                                 ('CHA', (24, 24, 25), 'atWBII', 0.1, 3e-8)])
def test_real_framework_solution(name, shape, version, p_bar, tol_dilute):
    """Methane, hard spheres + mean-field WDA + LJ: the product program built
    from the input files against the oracle's Anderson solve (the fixed
    point; chabazite's deep wells, -24 kJ/mol, make the reference's static
    Picard crawl, and at 10 bar its ideal-gas start overfills the cages so
    that the reference's first steps are non-finite: 0.1 bar there).

    The RIUE criterion bounds the integrated change, so in chabazite's most
    dilute corners (rho ~ 4e-8 of its maximum) the converged density is fixed
    only to ~1e-8 relative: the oracle started from rho0 (1 + 1e-12) lands
    1.15e-8 from its own unperturbed solution there, and 1.2e-9 where rho
    exceeds 1e-4 of the maximum.  Hence 3e-8 over every rho > 1e-10 A^-3 and
    1e-8 over the dense part."""
    from cmmdft_amd.program import Program
    from cmmdft_amd.system import SphericalLJGuest
    from cmmdft_amd.eos import MFAEOS
    from cmmdft_amd.solver import Anderson, grand_potential
    h = _host(name)
    og, parts, beta, lam, V = oracle_system(_as_synthetic(h), shape, version)
    mu = mu_at(p_bar)
    rho0 = O.initial_density(og, mu, beta, lam, V=V)
    s = O.Anderson(og, parts, beta, lam, nsteps=1000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                   threshold=1e-13)
    Nref, rho_ref, conv = s.solve(mu, rho0)
    assert conv
    omega_ref = s.omega(rho_ref, og.fft(rho_ref))
    prog = Program(write_files=False)
    prog.set_system(h, SphericalLJGuest('CH4', W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    prog.set_grid(npoints=list(shape))
    prog.init_free_energy(T)
    prog.fener.add_hard_sphere(version)
    prog.fener.add_wdav(MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    prog.fener.add_external_potential(cutoff=W.CUTOFF)
    prog.set_solver(Anderson(prog, nsteps=1000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                             threshold=1e-13))
    N, rho, conv = prog.solve(mu)
    assert conv
    rel, mism = rho_rel_err(rho, rho_ref)
    assert rel <= tol_dilute, rel
    dense = rho_ref > 1e-4 * rho_ref.max()
    assert np.max(np.abs(rho[dense] - rho_ref[dense]) / rho_ref[dense]) <= 1e-8
    assert abs(N - Nref) <= 1e-8 * Nref
    om = grand_potential(prog.fener, mu, rho)
    assert abs(om - omega_ref) <= 1e-8 * abs(omega_ref), (om, omega_ref)


@pytest.mark.parametrize('name,shape', [('ZIF8', (30, 32, 28)), ('CHA', (28, 28, 30))])
def test_distance_and_regions(name, shape):
    """Program.calc_distance (device nearest-atom distance, fractional minimum
    image) and calc_regions against the oracle (program.py:145-217)."""
    from cmmdft_amd.program import Program
    from cmmdft_amd.system import SphericalLJGuest
    h = _host(name)
    prog = Program(write_files=False)
    prog.set_system(h, SphericalLJGuest('CH4', W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    prog.set_grid(npoints=list(shape))
    prog.init_free_energy(T)
    prog.fener.add_external_potential(cutoff=W.CUTOFF)
    site, mof, empty = prog.calc_regions()
    og = O.Grid(np.asarray(h.cell.rvecs), shape)
    dref = O.min_distance(og.points[..., :3], h.mol.pos, np.asarray(h.cell.rvecs)).reshape(shape)
    assert np.abs(prog.dis - dref).max() <= 1e-12 * dref.max()
    ep = prog.fener.part_dict['ExtPot'].potential
    rs, rm, re = O.regions(dref, ep, T)
    assert np.array_equal(site, rs) and np.array_equal(mof, rm) and np.array_equal(empty, re)
    assert site.any() and mof.any() and empty.any()
    c = prog.copy()
    assert c.grid is prog.grid and c.fener is prog.fener and c.system is prog.system
