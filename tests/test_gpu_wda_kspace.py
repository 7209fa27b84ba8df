"""The linear WDA in k space.  With every EOS term mean field the WDA's
excess free energy per volume is f_ex = A rho_bar^2 (eos.py MFAEOS), so its
derivative w_D * 2A rho_bar is linear in rho; the device applies it as
2A w_D^2 rho^ in the x passes and takes its energy by Parseval, without
forming rho_bar.  The pointwise route (WDAVFunctional(kspace=False), the one
every other EOS takes) computes the same quantities in real space.  The two
must agree to rounding on the same input -- derivative, energy sums, Hessian
product, and a converged solve -- and the k-space route must be the one the
plan takes for mean-field EOS terms only.  (The oracle parity of the k-space
route is covered where the default route runs: the WDA and solver parity
tests.)"""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from helpers import T, mu_at, product_program

pytestmark = pytest.mark.gpu

SHAPE = (32, 32, 32)


@pytest.fixture(scope='module')
def prog():
    return product_program(W.host_case2(), SHAPE, 'WBII')


def _density(prog, seed=7):
    import torch
    rho = prog._set_initial_density(chempot=mu_at(10.0), temperature=T).reshape(SHAPE)
    g = torch.Generator(device='cpu').manual_seed(seed)
    return (rho * (1 + 0.3 * torch.rand(SHAPE, generator=g, dtype=torch.float64).cuda())).contiguous()


def _chain(prog, kspace, rho):
    import torch
    from cmmdft_amd import _lib as L
    from cmmdft_amd.solver import _Engine
    prog.fener.part_dict['WDA-V'].kspace = kspace
    try:
        eng = _Engine(prog.grid, prog.fener)
        assert eng.plan.info['wda_kspace'] == (1 if kspace else 0)
        r = rho.reshape(-1).contiguous()
        c = torch.empty_like(r)
        eng.plan.seed(r, None, False, eng.lam3)
        eng.plan.derive()
        eng.plan.finish_c1(r, None, False, c, eng.lam3)
        s = eng.plan.scalars()[0].copy()
        eng.plan.seed(r, None, False, eng.lam3)
        eng.plan.value()
        s2 = eng.plan.scalars()[0].copy()
        return c.cpu().numpy(), s, s2
    finally:
        prog.fener.part_dict['WDA-V'].kspace = True


def test_derivative_and_energies_agree(prog):
    from cmmdft_amd import _lib as L
    rho = _density(prog)
    ck, sk, vk = _chain(prog, True, rho)
    cp, sp, vp = _chain(prog, False, rho)
    scale = np.abs(cp).max()
    assert np.abs(ck - cp).max() <= 1e-12 * scale, np.abs(ck - cp).max() / scale
    for s_k, s_p in ((sk, sp), (vk, vp)):
        assert abs(s_k[L.S_FEX] - s_p[L.S_FEX]) <= 1e-12 * abs(s_p[L.S_FEX])
        assert s_k[L.S_PHI] == pytest.approx(s_p[L.S_PHI], rel=1e-13)
        assert s_k[L.S_N3MAX] == s_p[L.S_N3MAX]
    # derive and value report the same WDA energy on either route
    assert vk[L.S_FEX] == pytest.approx(sk[L.S_FEX], rel=1e-14)


def test_wda_alone_and_energy_by_parseval(prog):
    """The WDA part alone (no hard spheres): derivative and value on both
    routes, and the energy against A sum rho_bar^2 dr from the weighted
    density the full chain forms."""
    import torch
    w = prog.fener.part_dict['WDA-V']
    rho = _density(prog, seed=11)
    out = {}
    for ks in (True, False):
        w.kspace = ks
        try:
            out[ks] = (w.derive(rho).cpu().numpy(), w.value(rho))
        finally:
            w.kspace = True
    scale = np.abs(out[False][0]).max()
    assert np.abs(out[True][0] - out[False][0]).max() <= 1e-12 * scale
    assert out[True][1] == pytest.approx(out[False][1], rel=1e-12)
    rb = w._get_weighted_density(rho).cpu().numpy()
    A = w.eos.device_terms()[0][1] * w.eos.device_terms()[0][2][0]
    assert out[True][1] == pytest.approx(A * np.sum(rb * rb) * prog.grid.dr, rel=1e-12)


def test_hessian_product_agrees(prog):
    import torch
    from cmmdft_amd.adjoint import AdjointSolver
    rho = _density(prog, seed=3)
    g = torch.Generator(device='cpu').manual_seed(5)
    v = (torch.rand(SHAPE, generator=g, dtype=torch.float64) - 0.5).cuda() * 1e-4 / U.angstrom ** 3
    w = prog.fener.part_dict['WDA-V']
    out = {}
    for ks in (True, False):
        w.kspace = ks
        try:
            out[ks] = AdjointSolver(prog).hessian_vector(rho, v).cpu().numpy()
        finally:
            w.kspace = True
    scale = np.abs(out[False]).max()
    assert np.abs(out[True] - out[False]).max() <= 1e-12 * scale


@pytest.mark.parametrize('n3_series', [1e-8, 1e-5])
def test_solve_agrees(prog, n3_series):
    """Converged solutions (RIUE < 1e-13) of both routes.  With the
    reference's White-Bear II forms (series below n3 = 1e-8) the closed forms
    of phi2, phi3 lose their digits for 1e-8 < n3 < 1e-5, so at points of
    such packing a rounding-level change of the input moves the derivative
    by far more than rounding, and the solutions agree there only to that
    level; with the series up to 1e-5 they agree to 1e-9."""
    from cmmdft_amd.solver import Anderson
    mu = mu_at(10.0)
    w = prog.fener.part_dict['WDA-V']
    hs = prog.fener.part_dict['HardSphere']
    res = {}
    hs.n3_series = n3_series
    try:
        for ks in (True, False):
            w.kspace = ks
            s = Anderson(prog, nsteps=400, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-13)
            N, rho, conv = s.solve(mu, prog._set_initial_density(chempot=mu, temperature=T))
            assert conv
            res[ks] = (N, rho.cpu().numpy())
    finally:
        w.kspace = True
        hs.n3_series = 1e-8
    assert abs(res[True][0] - res[False][0]) <= 1e-10 * res[False][0]
    m = res[False][1] > 1e-10 / U.angstrom ** 3
    rel = np.abs(res[True][1][m] - res[False][1][m]) / res[False][1][m]
    i = int(np.argmax(rel))
    bound = 1e-9 if n3_series >= 1e-5 else 1e-5
    assert rel[i] <= bound, (rel[i], res[False][1][m][i] * U.angstrom ** 3)


@pytest.mark.parametrize('terms,linear', [('mfa2', True), ('mfa+cs', False)])
def test_route_follows_the_eos_terms(terms, linear):
    """Sums of mean-field terms stay linear (k space); a Carnahan-Starling
    term sends the WDA down the pointwise route."""
    from cmmdft_amd import eos as E
    from cmmdft_amd.functionals import WDAVFunctional, build_functional
    from cmmdft_amd.system import Grid
    m, s, e = W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON
    R = 1.8 * U.angstrom
    mfa = E.MFAEOS(m, s, e)
    eos = (E.SumOfEOS(m, [mfa, E.MFAEOS(m, 0.9 * s, 1.1 * e)], [1, -0.5]) if terms == 'mfa2'
           else E.SumOfEOS(m, [mfa, E.CarnahanStarlingEOS(m, R)], [1, 1]))
    g = Grid(W.host_case1().rvecs, (16, 16, 16))
    w = WDAVFunctional(g, R, eos)
    w.set_temperature(T, R)
    plan = g.plan()
    plan.set_functional(build_functional(wda=w, kT=U.boltzmann * T))
    assert plan.info['wda_kspace'] == (1 if linear else 0)
