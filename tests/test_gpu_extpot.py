"""Device LJ grid sum against the oracle's restatement of
get_external_potential (cmmdft/extpot_calculator.py:186-222)."""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W
from oracle import cdft_oracle as O

pytestmark = pytest.mark.gpu


def _rel_err(V, Vr):
    # tolerance metric: max |dV| / max |V|, and pointwise relative error where
    # |V| exceeds 1e-6 of the maximum
    scale = np.abs(Vr).max()
    big = np.abs(Vr) > 1e-6 * scale
    return np.abs(V - Vr).max() / scale, (np.abs(V - Vr)[big] / np.abs(Vr[big])).max()


@pytest.mark.parametrize('case,shape', [('case1', (48, 48, 48)), ('case1', (30, 36, 40)), ('case2', (64, 64, 64))])
def test_lj_grid_matches_oracle(case, shape):
    from cmmdft_amd.system import Grid
    from cmmdft_amd.extpot import get_external_potential
    host = W.host_case1() if case == 'case1' else W.host_case2()
    og = O.Grid(host.rvecs, shape)
    sig, eps = host.atom_params()
    Vr = O.external_potential(og.points[..., :3], sig, eps, host.pos, host.rvecs, W.GUEST_SIGMA, W.GUEST_EPSILON)
    V = get_external_potential(Grid(host.rvecs, shape), host.ff, W.GUEST_SIGMA, W.GUEST_EPSILON, host)
    a, r = _rel_err(V, Vr)
    assert a <= 1e-10 and r <= 1e-10, (a, r)
    # explicit point list form
    V2 = get_external_potential(og.points[..., :3].reshape(-1, 3), host.ff, W.GUEST_SIGMA, W.GUEST_EPSILON, host)
    a, r = _rel_err(V2, Vr)
    assert a <= 1e-10 and r <= 1e-10, (a, r)


def test_lj_single_atom_analytic():
    """One atom at the origin: V(r) = 4 eps ((s/R)^12 - (s/R)^6) - shift inside r_c."""
    from cmmdft_amd.system import Grid
    from cmmdft_amd.extpot import get_external_potential
    a = 30.0 * U.angstrom
    host = W.SyntheticHost(np.eye(3) * a, np.zeros((1, 3)), np.array([0]), ['X'],
                           {'X': (3.2 * U.angstrom, 40 * U.boltzmann)})
    g = Grid(host.rvecs, (40, 40, 40))
    V = get_external_potential(g, host.ff, W.GUEST_SIGMA, W.GUEST_EPSILON, host)
    p = g.points[..., :3].reshape(-1, 3)
    L = a
    d = p - L * np.round(p / L)
    R = np.sqrt((d ** 2).sum(1) + 1e-16)
    s = 0.5 * (3.2 * U.angstrom + W.GUEST_SIGMA)
    e = np.sqrt(40 * U.boltzmann * W.GUEST_EPSILON)
    rc = 12 * U.angstrom
    ref = np.where(R < rc, 4 * e * ((s / R) ** 12 - (s / R) ** 6) - 4 * e * ((s / rc) ** 12 - (s / rc) ** 6), 0.0)
    assert np.abs(V - ref).max() <= 1e-12 * np.abs(ref).max()
