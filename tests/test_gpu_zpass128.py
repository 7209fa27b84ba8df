"""The functional passes with their tiles fixed at compile time for 128-point
axes: the z pass (k_zphi128: 2, 4, 6 or 8 pencils per block), the one every
128-point z axis takes, and the x passes (k_xweights / k_xback with 2 or 4
columns, nx = 128, radix sequence 8, 4, 4; 1 or 3 columns take the runtime
tile) and the y pass (k_ypass128, 8, 13 or 16 columns at ny = 128 with that
radix sequence).  The derivative is the same bit for bit for every tile (each
pencil's and column's transforms and pointwise stage do the same operations
whatever the tile), the energy sums agree to rounding (their block partials
differ), and at 128^3 the hard-sphere derivative and value match the
oracle's restatement of HardSphereFunctional
(cmmdft/functionals.py:379-618)."""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O
from helpers import T, mu_at, product_program

pytestmark = pytest.mark.gpu


def _derive_at_tiles(shape, kspace, tiles, mfa=False):
    import torch
    from cmmdft_amd.solver import _Engine
    prog = product_program(W.host_case2(), shape, 'WBII')
    prog.fener.part_dict['WDA-V'].kspace = kspace
    if mfa:
        prog.fener.add_mean_field()
    eng = _Engine(prog.grid, prog.fener)
    assert eng.plan.info['wda_kspace'] == int(kspace)
    if shape[0] == 128:
        eng.plan.set_radix(0, [8, 4, 4])
    if shape[1] == 128:
        eng.plan.set_radix(1, [8, 4, 4])
    rho = prog._set_initial_density(chempot=mu_at(10.0), temperature=T).reshape(-1).contiguous()
    g = torch.Generator(device='cpu').manual_seed(2)
    rho = (rho * (1 + 0.3 * torch.rand(rho.shape, generator=g, dtype=torch.float64).cuda())).contiguous()
    out = {}
    for key, kw in tiles:
        eng.plan.set_tiles(**kw)
        c = torch.empty_like(rho)
        eng.plan.seed(rho, None, False, eng.lam3)
        eng.plan.derive()
        eng.plan.finish_c1(rho, None, False, c, eng.lam3)
        out[key] = (c.clone(), eng.plan.scalars()[0].copy())
    return out


def _same(out, ref):
    import torch
    from cmmdft_amd import _lib as L
    for key in out:
        assert torch.equal(out[key][0], out[ref][0]), key
        for q in (L.S_PHI, L.S_FEX, L.S_FMFA):
            assert out[key][1][q] == pytest.approx(out[ref][1][q], rel=1e-13, abs=1e-300)
        assert out[key][1][L.S_N3MAX] == out[ref][1][L.S_N3MAX]


@pytest.mark.parametrize('shape', [(128, 128, 128), (16, 20, 128), (10, 25, 128)])
@pytest.mark.parametrize('kspace', [True, False])
def test_tiles_give_the_same_derivative(shape, kspace):
    """Every pencil count gives the same bits.  The grids' last tiles are
    partial for P = 6 at 16 x 20 and for P = 4, 6 and 8 at 10 x 25 (250
    pencils), where the loader clamps and zero-fills; P = 2 tiles are full."""
    out = _derive_at_tiles(shape, kspace, [(P, dict(p_phi=P)) for P in (2, 4, 6, 8)])
    _same(out, 6)


@pytest.mark.parametrize('kspace', [True, False])
def test_runtime_z_tiles_give_the_same_derivative(kspace):
    """The runtime-shaped z pass (nz != 128), which loads and stores its tile
    by (pencil pair, kz) columns like the compile-time one: every pencil
    count gives the same bits, with partial last tiles at 10 x 25 (250
    pencils) for P = 4, 6 and 8 and full ones for P = 2 and 10."""
    out = _derive_at_tiles((10, 25, 24), kspace, [(P, dict(p_phi=P)) for P in (2, 4, 6, 8, 10)])
    _same(out, 2)


@pytest.mark.parametrize('shape', [(128, 128, 128), (128, 24, 20)])
@pytest.mark.parametrize('kspace,mfa', [(True, False), (False, False), (True, True)])
def test_x_tiles_give_the_same_derivative(shape, kspace, mfa):
    """2 and 4 columns (compile-time tiles) against 1 and 3 (runtime)."""
    out = _derive_at_tiles(shape, kspace, [(tb, dict(tb_x=tb)) for tb in (1, 2, 3, 4)], mfa=mfa)
    _same(out, 1)


@pytest.mark.parametrize('shape', [(128, 128, 128), (20, 128, 24)])
@pytest.mark.parametrize('kspace,mfa', [(True, False), (False, False), (True, True)])
def test_y_tiles_give_the_same_derivative(shape, kspace, mfa):
    """8, 13 and 16 columns (the compile-time y pass at ny = 128, whose
    paired block keeps its tile in registers) against 5 and 12 (the runtime
    y pass, whose paired block works in rounds of columns)."""
    out = _derive_at_tiles(shape, kspace, [(tb, dict(tb_y=tb)) for tb in (5, 8, 12, 13, 16)], mfa=mfa)
    _same(out, 12)


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
