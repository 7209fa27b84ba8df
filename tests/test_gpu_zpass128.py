"""The functional z pass for nz = 128 with its tile fixed at compile time
(k_zphi128: 2, 4, 6 or 8 pencils per block), the one every 128-point z axis
takes: the derivative is the same bit for bit for every tile (each pencil's
transforms and pointwise stage do the same operations whatever the tile),
the energy sums agree to rounding (their block partials differ), and at
128^3 the hard-sphere derivative and value match the oracle's restatement
of HardSphereFunctional (cmmdft/functionals.py:379-618)."""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, product_program

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('shape', [(128, 128, 128), (16, 20, 128)])
@pytest.mark.parametrize('kspace', [True, False])
def test_tiles_give_the_same_derivative(shape, kspace):
    import torch
    from cmmdft_amd import _lib as L
    from cmmdft_amd.solver import _Engine
    prog = product_program(W.host_case2(), shape, 'WBII')
    prog.fener.part_dict['WDA-V'].kspace = kspace
    eng = _Engine(prog.grid, prog.fener)
    assert eng.plan.info['wda_kspace'] == int(kspace)
    rho = prog._set_initial_density(chempot=mu_at(10.0), temperature=T).reshape(-1).contiguous()
    g = torch.Generator(device='cpu').manual_seed(2)
    rho = (rho * (1 + 0.3 * torch.rand(rho.shape, generator=g, dtype=torch.float64).cuda())).contiguous()
    out = {}
    for P in (2, 4, 6, 8):
        eng.plan.set_tiles(p_phi=P)
        c = torch.empty_like(rho)
        eng.plan.seed(rho, None, False, eng.lam3)
        eng.plan.derive()
        eng.plan.finish_c1(rho, None, False, c, eng.lam3)
        out[P] = (c.clone(), eng.plan.scalars()[0].copy())
    for P in (2, 4, 8):
        assert torch.equal(out[P][0], out[6][0]), P
        for q in (L.S_PHI, L.S_FEX):
            assert out[P][1][q] == pytest.approx(out[6][1][q], rel=1e-13)
        assert out[P][1][L.S_N3MAX] == out[6][1][L.S_N3MAX]


@pytest.mark.parametrize('version', ['WBII', 'aFMT'])
def test_hard_sphere_at_128_cube_matches_oracle(version):
    import torch
    from cmmdft_amd.functionals import HardSphereFunctional
    from cmmdft_amd.system import Grid
    host = W.host_case2()
    shape = (128, 128, 128)
    og = O.Grid(host.rvecs, shape)
    sig, eps = host.atom_params()
    V = O.external_potential(og.points[..., :3], sig, eps, host.pos, host.rvecs, W.GUEST_SIGMA,
                             W.GUEST_EPSILON).reshape(shape)
    beta = 1 / (U.boltzmann * T)
    rng = np.random.default_rng(4)
    rho = 3e-3 / U.angstrom ** 3 * np.exp(-beta * np.clip(V, -50 * U.kjmol, 50 * U.kjmol)) * \
        (1 + 0.2 * rng.random(shape))
    R = O.bh_radius(W.GUEST_SIGMA, W.GUEST_EPSILON, T)
    ohs = O.HardSphere(og, R, version, T)
    krho = og.fft(rho)
    c_ref = ohs.derive(krho)
    F_ref = ohs.value(krho)
    hs = HardSphereFunctional(Grid(host.rvecs, shape), R, version)
    hs.set_temperature(T, R)
    c = hs.derive(torch.from_numpy(rho).cuda()).cpu().numpy()
    # aFMT: bound relative to max|c|; WBII: on beta |dc| (its phi2/phi3
    # closed forms cancel at low n3, so rounding-level input changes move c
    # by ~1e-9 relative there)
    tol = 1e-12 * np.abs(c_ref).max() if version == 'aFMT' else 1e-8 * U.boltzmann * T
    assert np.abs(c - c_ref).max() <= tol
    F = hs.value(torch.from_numpy(rho).cuda())
    assert abs(F - F_ref) <= 1e-11 * abs(F_ref)
