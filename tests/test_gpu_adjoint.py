"""Linear response of the converged density (adjoint): the dual-number
Hessian product against finite differences of the HIP functional derivative,
the LJ backward kernel against finite differences of the LJ grid sum, and
dN/dmu, dN/dsigma_t, dN/depsilon_t against central finite differences of full
solves (relative step 1e-4), on the GPU and with the oracle on the CPU."""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, oracle_system, product_program

pytestmark = pytest.mark.gpu

SHAPE = (24, 24, 24)


def _structured_density(grid, host, seed=5):
    import torch
    from cmmdft_amd.extpot import get_external_potential
    V = get_external_potential(grid, host.ff, W.GUEST_SIGMA, W.GUEST_EPSILON, host, cutoff=W.CUTOFF,
                               as_tensor=True).reshape(grid.npoints)
    beta = 1 / (U.boltzmann * T)
    rb = 3e-3 / U.angstrom ** 3
    rho = rb * torch.exp(-beta * torch.clamp(V, -50 * U.kjmol, 50 * U.kjmol))
    g = torch.Generator(device='cpu').manual_seed(seed)
    rho = rho * (1 + 0.2 * torch.rand(grid.npoints, generator=g, dtype=torch.float64).cuda())
    return rho


def _program(version, eos='MFA', host=None, shape=SHAPE, n3_series=1e-8):
    from cmmdft_amd.program import Program
    from cmmdft_amd.system import NanoporousHost, SphericalLJGuest
    from cmmdft_amd import eos as E
    host = host or W.host_case4()
    prog = Program(write_files=False)
    prog.set_system(NanoporousHost('host', host, host.ff),
                    SphericalLJGuest('CH4', W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    prog.set_grid(npoints=list(shape))
    prog.init_free_energy(T)
    if version:
        prog.fener.add_hard_sphere(version, n3_series=n3_series)
    if eos == 'MFA':
        prog.fener.add_wdav(E.MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON))
    elif eos == 'corr':
        m = E.ModifiedBenedictWebbRubinEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON)
        cs = E.CarnahanStarlingEOS(W.GUEST_MASS, prog.system.guest.Rhs)
        mfa = E.MFAEOS(W.GUEST_MASS, W.GUEST_SIGMA, W.GUEST_EPSILON)
        prog.fener.add_wdav(E.SumOfEOS(W.GUEST_MASS, [m, cs, mfa], [1.0, -1.0, -1.0]))
    prog.fener.add_external_potential(cutoff=W.CUTOFF)
    return prog, host


def _c_ex(prog, rho):
    """c_ex(rho) through the HIP functional chain (no external potential)."""
    import torch
    from cmmdft_amd.solver import _Engine
    eng = _Engine(prog.grid, prog.fener)
    r = rho.reshape(-1).contiguous()
    out = torch.empty_like(r)
    eng.plan.seed(r, None, False, eng.lam3)
    eng.plan.derive()
    eng.plan.finish_c1(r, None, False, out, eng.lam3)
    return out.reshape(prog.grid.npoints)


def _dense_density(grid, seed=7):
    """A fluid-like density 1.5e-3..4.5e-3 A^-3 everywhere (n3 ~ 0.05..0.1),
    where every branch of the pointwise forms is well conditioned."""
    import torch
    g = torch.Generator(device='cpu').manual_seed(seed)
    u = torch.rand(grid.npoints, generator=g, dtype=torch.float64).cuda()
    return (3e-3 / U.angstrom ** 3) * (1 + (u - 0.5))


def _check_hvp(prog, rho, seed, tol):
    import torch
    from cmmdft_amd.adjoint import AdjointSolver
    g = torch.Generator(device='cpu').manual_seed(seed)
    v = (torch.rand(prog.grid.npoints, generator=g, dtype=torch.float64).cuda() - 0.5) * rho
    Mv = AdjointSolver(prog).hessian_vector(rho, v)
    # 4th-order central differences of the HIP c_ex along v
    h = 1e-3
    c = [_c_ex(prog, rho + k * h * v) for k in (-2, -1, 1, 2)]
    fd = (c[0] - 8 * c[1] + 8 * c[2] - c[3]) / (12 * h)
    err = (Mv - fd).abs().max().item()
    scale = fd.abs().max().item()
    assert err <= tol * scale, (err, scale)


@pytest.mark.parametrize('version,eos', [('WBII', 'MFA'), ('FMT', 'MFA'), ('aWBII', 'MFA'), ('MFMT', None),
                                         (None, 'corr'), ('aFMT', 'corr'), ('WBII', 'corr'), ('tFMT', 'MFA'),
                                         ('atWBII', 'MFA'), ('tMFMT', 'corr')])
def test_hessian_vector_matches_finite_differences(version, eos):
    prog, host = _program(version, eos)
    _check_hvp(prog, _dense_density(prog.grid), 11, 1e-8)


@pytest.mark.parametrize('version,eos', [('FMT', 'MFA'), ('aFMT', 'corr'), (None, 'corr'), ('tFMT', 'MFA')])
def test_hessian_vector_near_empty_regions(version, eos):
    """Densities ~1e-12 near the host atoms (n3 down to the clip).  The White-
    Bear forms are left out here: at n3 << 1 they lose ~1e-9 of relative
    precision to cancellation, which the finite differences amplify."""
    prog, host = _program(version, eos)
    _check_hvp(prog, _structured_density(prog.grid, host), 13, 1e-8)


@pytest.mark.parametrize('version,eos', [('WBII', 'corr'), ('FMT', 'MFA'), ('aWBII', 'MFA'), ('atWBII', 'MFA'),
                                         ('tFMT', 'corr')])
def test_hessian_is_symmetric(version, eos):
    """u.Mv = v.Mu.  (Where the reference's clips act -- n3 below 1e-30, or
    xi above 1 for the 'a' forms -- its c_ex is not the gradient of F_ex and M
    is not symmetric; that happens only where rho ~ 0, which the adjoint
    weights by sqrt(beta rho).  The density here has no clipped points.)"""
    import torch
    from cmmdft_amd.adjoint import AdjointSolver
    prog, host = _program(version, eos)
    rho = _dense_density(prog.grid)
    g = torch.Generator(device='cpu').manual_seed(12)
    u = (torch.rand(prog.grid.npoints, generator=g, dtype=torch.float64).cuda() - 0.5) * rho
    v = (torch.rand(prog.grid.npoints, generator=g, dtype=torch.float64).cuda() - 0.5) * rho
    adj = AdjointSolver(prog)
    Mv = adj.hessian_vector(rho, v)
    uMv = float((u * Mv).sum())
    vMu = float((v * adj.hessian_vector(rho, u)).sum())
    bound = 1e-12 * float(u.norm() * Mv.norm())
    assert abs(uMv - vMu) <= bound, (uMv, vMu, bound)


@pytest.mark.parametrize('images,limit', [(0, 0.0), (0, 5 * U.kjmol), (1, 0.0)])
def test_lj_backward_matches_finite_differences(images, limit):
    import torch
    from cmmdft_amd import _lib as L
    from cmmdft_amd.extpot import host_atom_table
    host = W.host_case4()
    shape = (20, 18, 22)
    rv = host.rvecs
    types = torch.from_numpy(np.asarray(host.ffatype_ids, dtype=np.int32)).cuda()
    g = torch.Generator(device='cpu').manual_seed(3)
    w = torch.rand(int(np.prod(shape)), generator=g, dtype=torch.float64).cuda()

    def V_of(ff):
        tab = torch.from_numpy(host_atom_table(ff, host)).cuda()
        out = torch.empty(int(np.prod(shape)), dtype=torch.float64, device='cuda')
        L.lj_grid(shape, rv, tab, tab.shape[0], W.GUEST_SIGMA, W.GUEST_EPSILON, W.CUTOFF, images, 0.0, out)
        return out

    tab = torch.from_numpy(host_atom_table(host.ff, host)).cuda()
    G = L.lj_grid_grad(shape, rv, tab, types, tab.shape[0], 4, W.GUEST_SIGMA, W.GUEST_EPSILON, W.CUTOFF, images,
                       limit, w).cpu().numpy()
    V0 = V_of(host.ff)
    clamped = (~torch.isfinite(V0) | (V0 > limit)) if limit > 0 else torch.zeros_like(V0, dtype=torch.bool)
    if limit > 0:
        assert 0 < int(clamped.sum()) < V0.numel()
    for t, name in enumerate(host.ffatypes):
        for j in range(2):
            h = 1e-4 * host.ff[name][j]
            vals = []
            for k in (-2, -1, 1, 2):
                ff = dict(host.ff)
                par = list(ff[name])
                par[j] += k * h
                ff[name] = tuple(par)
                Vk = V_of(ff)
                # the kernel's derivative is zero where the potential is clamped
                vals.append(float((w * torch.where(clamped, torch.zeros_like(Vk), Vk)).sum()))
            fd = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
            assert abs(G[t, j] - fd) <= 1e-7 * abs(fd), (t, j, G[t, j], fd)


def _solve(prog, mu, threshold=1e-13):
    from cmmdft_amd.solver import Anderson
    prog.set_solver(Anderson(prog, nsteps=2000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                             threshold=threshold))
    N, rho, conv = prog.solve(mu)
    assert conv
    return N, rho


def test_ideal_gas_closed_form():
    """Without an excess functional J = diag(kT/rho): lambda = beta rho dr,
    dN/dmu = beta N, dN/dtheta = -beta sum rho dV/dtheta dr."""
    import torch
    from cmmdft_amd.adjoint import AdjointSolver
    from cmmdft_amd import _lib as L
    from cmmdft_amd.extpot import host_atom_table
    prog, host = _program(None, None)
    mu = mu_at(10.0)
    N, rho = _solve(prog, mu)
    adj = AdjointSolver(prog, tol=1e-13)
    g = adj.gradients(rho)
    beta = prog.fener.beta
    assert adj.iters <= 2
    assert abs(g['dN_dmu'] - beta * N) <= 1e-11 * beta * N
    r = torch.from_numpy(rho).cuda().reshape(-1)
    tab = torch.from_numpy(host_atom_table(prog.system.host.ff, prog.system.host.mol)).cuda()
    types = torch.from_numpy(np.asarray(host.ffatype_ids, dtype=np.int32)).cuda()
    ep = prog.fener.part_dict['ExtPot']
    Gd = L.lj_grid_grad(SHAPE, host.rvecs, tab, types, tab.shape[0], 4, W.GUEST_SIGMA, W.GUEST_EPSILON, W.CUTOFF, 0,
                        ep.limit_potential, r).cpu().numpy()
    ref = -beta * Gd * prog.grid.dr
    np.testing.assert_allclose(g['dN_dsigma'], ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g['dN_depsilon'], ref[:, 1], rtol=1e-10)


def _adjoint_case(version, n3_series=1e-8):
    from cmmdft_amd.adjoint import AdjointSolver, host_parameters
    prog, host = _program(version, 'MFA', n3_series=n3_series)
    mu = mu_at(10.0)
    N, rho = _solve(prog, mu)
    adj = AdjointSolver(prog, tol=1e-12)
    g = adj.gradients(rho)
    sig, eps = host_parameters(prog)
    return prog, host, mu, N, g, sig, eps


@pytest.fixture(scope='module')
def adjoint_case():
    return _adjoint_case('WBII')


@pytest.fixture(scope='module')
def adjoint_case_wb_series():
    return _adjoint_case('WBII', n3_series=1e-5)


@pytest.fixture(scope='module')
def adjoint_case_fmt():
    return _adjoint_case('FMT')


@pytest.fixture(scope='module')
def adjoint_case_tensor():
    return _adjoint_case('tFMT')


def _fd_host(prog, mu, sig, eps, t, j, rel=1e-4):
    """Central differences of N in parameter j (0 sigma, 1 epsilon) of host
    type t with relative steps rel and 2 rel: (plain, Richardson)."""
    from cmmdft_amd.adjoint import set_host_parameters
    h = rel * (sig, eps)[j][t]
    Ns = {}
    try:
        for k in (2, 1, -1, -2):
            s2, e2 = sig.copy(), eps.copy()
            (s2 if j == 0 else e2)[t] += k * h
            set_host_parameters(prog, s2, e2)
            Ns[k] = _solve(prog, mu)[0]
    finally:
        set_host_parameters(prog, sig, eps)
    d1 = (Ns[1] - Ns[-1]) / (2 * h)
    d2 = (Ns[2] - Ns[-2]) / (4 * h)
    return d1, (4 * d1 - d2) / 3


@pytest.mark.parametrize('case', ['adjoint_case', 'adjoint_case_fmt', 'adjoint_case_tensor'])
def test_dN_dmu_matches_finite_differences(case, request):
    prog, host, mu, N, g, sig, eps = request.getfixturevalue(case)
    h = 1e-4 * abs(mu)
    Ns = {k: _solve(prog, mu + k * h)[0] for k in (2, 1, -1, -2)}
    d1 = (Ns[1] - Ns[-1]) / (2 * h)
    rich = (4 * d1 - (Ns[2] - Ns[-2]) / (4 * h)) / 3
    assert abs(g['dN_dmu'] - d1) <= 1e-5 * abs(d1), (g['dN_dmu'], d1)
    assert abs(g['dN_dmu'] - rich) <= 1e-7 * abs(rich), (g['dN_dmu'], rich)


def test_host_parameter_gradients_match_finite_differences(adjoint_case_fmt):
    """All 8 host parameters, Rosenfeld FMT + mean-field WDA: central
    differences with relative step 1e-4 and their Richardson extrapolation
    (steps 1e-4 and 2e-4)."""
    prog, host, mu, N, g, sig, eps = adjoint_case_fmt
    grad = np.stack([g['dN_dsigma'], g['dN_depsilon']], axis=1)
    for t in range(4):
        for j in range(2):
            d1, rich = _fd_host(prog, mu, sig, eps, t, j)
            assert abs(grad[t, j] - d1) <= 1e-5 * abs(d1), (t, j, grad[t, j], d1)
            assert abs(grad[t, j] - rich) <= 2e-6 * abs(rich), (t, j, grad[t, j], rich)


def test_host_parameter_gradients_tensor(adjoint_case_tensor):
    """Tensor Rosenfeld FMT ('tFMT') + mean-field WDA: epsilon and sigma of two
    host types against central differences (relative step 1e-4)."""
    prog, host, mu, N, g, sig, eps = adjoint_case_tensor
    for t in (0, 2):
        for j in range(2):
            d1, rich = _fd_host(prog, mu, sig, eps, t, j)
            grad = (g['dN_dsigma'], g['dN_depsilon'])[j][t]
            assert abs(grad - d1) <= 1e-5 * abs(d1), (t, j, grad, d1)


def test_host_parameter_gradients_white_bear(adjoint_case, adjoint_case_wb_series):
    """White-Bear II, all 8 host parameters at 1e-5 against central
    differences (relative step 1e-4), with the series forms of phi2/phi3
    used up to n3 = 1e-5.  With the reference's switch at 1e-8 the closed
    forms lose digits for 1e-8 < n3 < 1e-5 (dphi3/dn3 keeps none for
    3e-8 < n3 < 1e-6; c1 moves by up to beta|dc| ~ 6e-8 at the 24^3 solution),
    so N(sigma) is non-smooth and the sigma differences scatter by up to
    ~2e-4 without a trend in the step, while the large-step (1e-3)
    differences still agree with the adjoint; epsilon, set by the attractive
    wells where n3 is large, is unaffected and checked there at 1e-5 too."""
    prog, host, mu, N, g, sig, eps = adjoint_case_wb_series
    for t in range(4):
        for j in range(2):
            d1, _ = _fd_host(prog, mu, sig, eps, t, j)
            grad = (g['dN_dsigma'], g['dN_depsilon'])[j][t]
            assert abs(grad - d1) <= 1e-5 * abs(d1), (t, j, grad, d1)
    prog, host, mu, N, g, sig, eps = adjoint_case
    for t in range(4):
        d1, _ = _fd_host(prog, mu, sig, eps, t, 1)
        assert abs(g['dN_depsilon'][t] - d1) <= 1e-5 * abs(d1), (t, g['dN_depsilon'][t], d1)
        d1, _ = _fd_host(prog, mu, sig, eps, t, 0, rel=1e-3)
        assert abs(g['dN_dsigma'][t] - d1) <= 3e-5 * abs(d1), (t, g['dN_dsigma'][t], d1)


def test_gradients_match_oracle_finite_differences(adjoint_case):
    """Reference-path finite differences: the oracle's own Anderson solves on
    the CPU, for dN/dmu and two host parameters."""
    prog, host, mu, N, g, sig, eps = adjoint_case
    og, parts, beta, lam, V = oracle_system(host, SHAPE, 'WBII')

    def oracle_N(mu_, sig_t, eps_t):
        Vx = O.external_potential(og.points[..., :3], sig_t[host.ffatype_ids], eps_t[host.ffatype_ids], host.pos,
                                  host.rvecs, W.GUEST_SIGMA, W.GUEST_EPSILON).reshape(og.npoints)
        ps = parts[:-1] + [O.ExtPot(og, Vx)]
        rho0 = O.initial_density(og, mu_, beta, lam, V=ps[-1].potential)
        a = O.Anderson(og, ps, beta, lam, nsteps=2000, method='anderson', m=5, damping=0.1, criterion='RIUE',
                       threshold=1e-13)
        Nx, _, conv = a.solve(mu_, rho0)
        assert conv
        return Nx

    assert abs(oracle_N(mu, sig, eps) - N) <= 1e-8 * N
    h = 1e-4 * abs(mu)
    fd = (oracle_N(mu + h, sig, eps) - oracle_N(mu - h, sig, eps)) / (2 * h)
    assert abs(g['dN_dmu'] - fd) <= 1e-5 * abs(fd), (g['dN_dmu'], fd)
    for t, j in ((0, 1), (2, 0)):
        base = (sig, eps)[j]
        h = 1e-4 * base[t]
        vals = []
        for k in (1, -1):
            s2, e2 = sig.copy(), eps.copy()
            (s2 if j == 0 else e2)[t] += k * h
            vals.append(oracle_N(mu, s2, e2))
        fd = (vals[0] - vals[1]) / (2 * h)
        grad = (g['dN_dsigma'], g['dN_depsilon'])[j][t]
        assert abs(grad - fd) <= 1e-5 * abs(fd), (t, j, grad, fd)


def test_autograd_uptake(adjoint_case):
    import torch
    from cmmdft_amd.adjoint import uptake
    prog, host, mu, N, g, sig, eps = adjoint_case
    from cmmdft_amd.solver import Anderson
    solver = Anderson(prog, nsteps=2000, method='anderson', m=5, damping=0.1, criterion='RIUE', threshold=1e-13)
    m = torch.tensor(mu, dtype=torch.float64, requires_grad=True)
    s = torch.tensor(sig, dtype=torch.float64, requires_grad=True)
    e = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    Nt = uptake(prog, m, s, e, solver=solver, adjoint_tol=1e-12)
    assert abs(float(Nt.detach()) - N) <= 1e-10 * N
    (2.0 * Nt).backward()
    assert abs(float(m.grad) - 2 * g['dN_dmu']) <= 1e-9 * abs(2 * g['dN_dmu'])
    np.testing.assert_allclose(s.grad.numpy(), 2 * g['dN_dsigma'], rtol=1e-9)
    np.testing.assert_allclose(e.grad.numpy(), 2 * g['dN_depsilon'], rtol=1e-9)


def test_config4_gradients_at_128_cube():
    """Config 4 at its stated size: the case-4 host on the 128^3 grid at
    10 bar, White-Bear II + mean-field WDA.  dN/dmu and the sigma and
    epsilon gradients of two host types against central differences of GPU
    solves with relative step 1e-4, at 1e-5."""
    from cmmdft_amd.adjoint import AdjointSolver, host_parameters
    prog, host = _program('WBII', 'MFA', shape=(128, 128, 128))
    mu = mu_at(10.0)
    N, rho = _solve(prog, mu)
    g = AdjointSolver(prog, tol=1e-12).gradients(rho)
    sig, eps = host_parameters(prog)
    h = 1e-4 * abs(mu)
    fd = (_solve(prog, mu + h)[0] - _solve(prog, mu - h)[0]) / (2 * h)
    assert abs(g['dN_dmu'] - fd) <= 1e-5 * abs(fd), (g['dN_dmu'], fd)
    for t in (0, 1):
        for j in (0, 1):
            d1, _ = _fd_host(prog, mu, sig, eps, t, j)
            grad = (g['dN_dsigma'], g['dN_depsilon'])[j][t]
            assert abs(grad - d1) <= 1e-5 * abs(d1), (t, j, grad, d1)
