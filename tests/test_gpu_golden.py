"""The HIP path against the committed oracle fixtures
(tests/golden/oracle_small.npz): FMT c1 and F for four versions, the WDA
derivative, the LJ grid, and a converged WBII + WDA + LJ solution at 10 bar,
on a 16 x 12 x 10 grid of the case-1 cell."""
import os

import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from helpers import T, product_program, rho_rel_err

pytestmark = pytest.mark.gpu

FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'oracle_small.npz'))
SHAPE = tuple(int(n) for n in FIX['shape'])
BETA = 1 / (U.boltzmann * T)


def _grid():
    from cmmdft_amd.system import Grid
    return Grid(FIX['rvecs'], SHAPE)


@pytest.mark.parametrize('version', ['FMT', 'WBII', 'aWBII', 'MFMT'])
def test_hard_sphere_against_fixture(version):
    import torch
    from cmmdft_amd.functionals import HardSphereFunctional
    R = float(FIX['R'])
    hs = HardSphereFunctional(_grid(), R, version)
    hs.set_temperature(T, R)
    rho = torch.from_numpy(FIX['rho']).cuda()
    c = hs.derive(rho).cpu().numpy()
    ref = FIX['c_' + version]
    if version == 'FMT':
        assert np.abs(c - ref).max() <= 1e-12 * np.abs(ref).max()
    else:   # the White-Bear forms' low-n3 conditioning: bound beta |dc| (see test_gpu_functional)
        assert BETA * np.abs(c - ref).max() <= 1e-8
    F = hs.value(rho)
    assert abs(F - float(FIX['F_' + version])) <= 1e-10 * abs(float(FIX['F_' + version]))


def test_wda_against_fixture():
    import torch
    from cmmdft_amd.functionals import WDAVFunctional
    from cmmdft_amd.eos import MFAEOS
    R = float(FIX['R'])
    wda = WDAVFunctional(_grid(), R, MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    wda.set_temperature(T, R)
    rho = torch.from_numpy(FIX['rho']).cuda()
    ref = FIX['c_WDA_MFA']
    assert np.abs(wda.derive(rho).cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(wda.value(rho) - float(FIX['F_WDA_MFA'])) <= 1e-12 * abs(float(FIX['F_WDA_MFA']))


def test_lj_grid_against_fixture():
    from cmmdft_amd.extpot import get_external_potential
    host = W.host_case1()
    V = get_external_potential(_grid(), host.ff, W.GUEST_SIGMA, W.GUEST_EPSILON, host, cutoff=W.CUTOFF)
    ref = FIX['V'].reshape(-1)
    m = np.abs(ref) < 1e4 * U.kjmol
    assert np.abs(V[m] - ref[m]).max() <= 1e-10 * np.abs(ref[m]).max()


@pytest.mark.parametrize('method', ['picard', 'anderson'])
def test_converged_solution_against_fixture(method):
    from cmmdft_amd.solver import Anderson, Picard
    from cmmdft_amd.solver import grand_potential
    host = W.host_case1()
    prog = product_program(host, SHAPE, 'WBII')
    if method == 'picard':
        prog.set_solver(Picard(prog, nsteps=20000, alpha_mix=0.05, method='static', criterion='RIUE',
                               threshold=1e-13))
    else:
        prog.set_solver(Anderson(prog, nsteps=2000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                                 threshold=1e-13))
    mu = float(FIX['mu'])
    N, rho, conv = prog.solve(mu)
    assert conv
    rel, mism = rho_rel_err(rho, FIX['rho_conv'])
    assert rel <= 1e-8
    if method == 'picard':
        assert mism == 0
    Nref = float(FIX['N_conv'])
    assert abs(N - Nref) <= 1e-8 * abs(Nref)
    om = grand_potential(prog.fener, mu, rho)
    oref = float(FIX['omega_conv'])
    assert abs(om - oref) <= 1e-8 * abs(oref)


EFF = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'oracle_effpot.npz'))
INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'inputs')


@pytest.mark.parametrize('name', ['co2', 'c2h6'])
def test_effective_potential_against_fixture(name):
    # tests/golden/oracle_effpot.npz: the oracle's float64 average (1e-10)
    # and the reference's float32 form (4e-7 relative, or 4e-7 kT)
    from cmmdft_amd.system import Grid
    from cmmdft_amd.extpot import get_external_potential_dict, generate_effective_potential
    host = W.host_case1()
    shape = tuple(int(n) for n in EFF['shape_%s' % name])
    gens = get_external_potential_dict(host.ff, os.path.join(INPUTS, '%s_pars_vdw_AA_trappe.txt' % name), host,
                                       cutoff=W.CUTOFF)
    V = generate_effective_potential(os.path.join(INPUTS, '%s_AA.chk' % name), Grid(EFF['rvecs'], shape), BETA, gens,
                                     degree=int(EFF['degree_%s' % name]))
    lim = 1e4 * U.kjmol

    def clamp(x):
        return np.minimum(np.where(np.isfinite(x), x, lim), lim)
    V, r64, r32 = clamp(V), clamp(EFF['effpot_%s_f64' % name]), clamp(EFF['effpot_%s_f32' % name])
    m = r64 < lim
    assert np.array_equal(V >= lim, r64 >= lim)
    assert np.abs(V[m] - r64[m]).max() <= 1e-10 * np.abs(r64[m]).max()
    assert np.all(np.abs(V - r32) <= 4e-7 * np.maximum(np.abs(r32), 1 / BETA))


def test_lj_derivative_table_and_interpolation_against_fixture():
    from cmmdft_amd.system import Grid
    from cmmdft_amd.extpot import get_external_potential_derivatives, Interpolator
    host = W.host_case1()
    tshape = tuple(int(n) for n in EFF['table_shape'])
    g = Grid(EFF['rvecs'], tshape)
    tab = get_external_potential_derivatives(g, host.ff, 3.05 * U.angstrom, 79.0 * U.boltzmann, host, g.spacings,
                                             cutoff=W.CUTOFF)
    for q in range(8):
        scale = np.abs(EFF['table'][q]).max()
        assert np.abs(tab[q] - EFF['table'][q]).max() <= 1e-10 * scale, q
    it = Interpolator(EFF['interp_table'], EFF['interp_origin'], EFF['interp_spacing'])
    pts = EFF['interp_points']
    scale = np.abs(EFF['interp_table']).max()
    for meth in ('tricubic', 'trilinear', 'tricubic_estimated'):
        assert np.abs(getattr(it, meth)(pts) - EFF[meth]).max() <= 1e-12 * 8 * scale, meth
