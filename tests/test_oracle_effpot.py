"""Pins of the oracle's orientational average for non-spherical guests
(cmmdft/extpot_calculator.py:348-389, 572-646) and of the product's host-side
quadrature tables; CPU only."""
import itertools
import math
import os

import numpy as np
import pytest

from cmmdft_amd import units as U
from cmmdft_amd.rotations import lebedev, combined_rotations, generate_rotation_matrix
from cmmdft_amd.system import ChkSystem
from cmmdft_amd.extpot import read_pars_file
from oracle import cdft_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, 'golden', 'inputs')
DEGREES = (3, 5, 7, 9, 11)


def _dfact(n):
    return 1 if n <= 0 else n * _dfact(n - 2)


@pytest.mark.parametrize('degree', DEGREES)
def test_lebedev_closed_forms_equal_reference_tables(degree):
    # the product's generated rules are the reference's stored tables, bit for bit
    p, w = lebedev(degree)
    with np.load(os.path.join(HERE, 'golden', 'lebedev_ref.npz')) as z:
        assert np.array_equal(p, z['points_%d' % degree])
        assert np.array_equal(w, z['weights_%d' % degree])
    # the irrational orbit coordinates are within an ulp of their closed forms
    vals = np.unique(np.abs(p))
    closed = {9: [np.sqrt((1 + 1 / np.sqrt(3)) / 2), np.sqrt((1 - 1 / np.sqrt(3)) / 2)],
              11: [1 / np.sqrt(11), 3 / np.sqrt(11)]}.get(degree, [])
    for c in closed:
        assert np.min(np.abs(vals - c)) <= 1.2e-16, c


@pytest.mark.parametrize('degree', DEGREES)
def test_lebedev_rules_integrate_monomials_exactly(degree):
    # sum_k w_k x^a y^b z^c = (a-1)!!(b-1)!!(c-1)!!/(a+b+c+1)!! (all even), 0 otherwise
    p, w = O.lebedev_table(degree)
    w = w / (4 * np.pi)
    for a, b, c in itertools.product(range(degree + 1), repeat=3):
        if a + b + c > degree:
            continue
        got = np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c)
        exact = 0.0
        if a % 2 == 0 and b % 2 == 0 and c % 2 == 0:
            exact = _dfact(a - 1) * _dfact(b - 1) * _dfact(c - 1) / _dfact(a + b + c + 1)
        assert abs(got - exact) <= 2e-14, (a, b, c, got, exact)


def test_lebedev_orbit_table_equals_reference_tables():
    """Every degree AngularGrid lists (3..131, AngGrid.py:38-71): the rule
    rebuilt from the orbit table has the SHA-256 of the reference's table
    (points and weights, float64), and equals the fixture tables of degree 13
    and 17 bit for bit."""
    import hashlib
    import json
    from cmmdft_amd.rotations import _ALL_DEGREES
    with open(os.path.join(HERE, 'golden', 'lebedev_digests.json')) as fh:
        dig = json.load(fh)
    assert sorted(int(d) for d in dig) == _ALL_DEGREES
    for d in _ALL_DEGREES:
        p, w = lebedev(d)
        assert len(p) == dig[str(d)]['size'], d
        h = hashlib.sha256(np.ascontiguousarray(p).tobytes() + np.ascontiguousarray(w).tobytes()).hexdigest()
        assert h == dig[str(d)]['sha256'], d
    with np.load(os.path.join(HERE, 'golden', 'lebedev_ref.npz')) as z:
        for d in (13, 17):
            p, w = lebedev(d)
            assert np.array_equal(p, z['points_%d' % d]) and np.array_equal(w, z['weights_%d' % d])
    # a requested degree maps to the next one up, as AngularGrid does
    assert len(lebedev(12)[0]) == 74 and len(lebedev(33)[0]) == 434
    with pytest.raises(ValueError):
        lebedev(132)


@pytest.mark.parametrize('degree', [13, 17, 23, 35, 53, 83, 131])
def test_lebedev_high_degree_rules_integrate_monomials_exactly(degree):
    """Exactness of the rebuilt rules for every monomial of total degree
    degree - 1 and degree, and for a sample of lower ones; points on the
    unit sphere and weights summing to 1."""
    p, w = lebedev(degree)
    assert np.allclose(np.sum(p * p, axis=1), 1.0, atol=1e-15)
    assert abs(w.sum() - 1.0) <= 1e-12
    n = degree
    px = np.stack([p[:, 0] ** k for k in range(n + 2)])
    py = np.stack([p[:, 1] ** k for k in range(n + 2)])
    pz = np.stack([p[:, 2] ** k for k in range(n + 2)])
    rng = np.random.default_rng(degree)
    tot = [n - 1, n] + [int(v) for v in rng.integers(0, n - 1, 4)]
    for t in tot:
        for a in range(t + 1):
            for b in range(t - a + 1):
                c = t - a - b
                got = np.sum(w * px[a] * py[b] * pz[c])
                exact = 0.0
                if a % 2 == 0 and b % 2 == 0 and c % 2 == 0:
                    exact = _dfact(a - 1) * _dfact(b - 1) * _dfact(c - 1) / _dfact(a + b + c + 1)
                assert abs(got - exact) <= 1e-13, (a, b, c, got, exact)
    # the low-degree rules are no more exact than stated: some even monomial
    # of degree + 1 has a visible error (high degrees make those integrals
    # too small to tell)
    if n <= 23:
        t = n + 1
        err = max(abs(np.sum(w * px[a] * py[b] * pz[t - a - b]) -
                      _dfact(a - 1) * _dfact(b - 1) * _dfact(t - a - b - 1) / _dfact(t + 1))
                  for a in range(0, t + 1, 2) for b in range(0, t - a + 1, 2))
        assert err > 1e-10, err


@pytest.mark.parametrize('degree', DEGREES)
def test_rotation_matrices(degree):
    R, w = O.rotation_matrix(degree, 3)
    pts, _ = O.lebedev_table(degree)
    eye = np.eye(3)
    for r in R:
        assert np.abs(r @ r.T - eye).max() <= 1e-15 * 4
        assert abs(np.linalg.det(r) - 1) <= 1e-14
    # R e_z is the Lebedev direction
    assert np.abs(R[:, :, 2] - pts).max() <= 4e-16
    R2, w2 = O.rotation_matrix(degree, 2)
    th = 2 * np.pi * np.arange(degree) / degree
    assert np.abs(R2[:, 0, 0] - np.cos(th)).max() <= 1e-15 and np.abs(R2[:, 1, 0] - np.sin(th)).max() <= 1e-15
    assert np.all(R2[:, 2, 2] == 1) and w2 == 1 / (degree * 4 * np.pi)
    # the product's matrices are the oracle's
    Rp, wp = generate_rotation_matrix(degree, 3)
    assert np.array_equal(Rp, R) and np.array_equal(wp, w)


@pytest.mark.parametrize('degree', (5, 9))
def test_combined_orientations_scale_columns(degree):
    # einsum('aij,bjk->abij') = R1_a diag(cos - sin, sin + cos, 1) for z angle b
    rot, w = O.orientations(degree)
    R1, w1 = O.rotation_matrix(degree, 3)
    th = 2 * np.pi * np.arange(degree) / degree
    for a in range(len(R1)):
        for b in range(degree):
            D = np.diag([np.cos(th[b]) - np.sin(th[b]), np.sin(th[b]) + np.cos(th[b]), 1.0])
            assert np.abs(rot[a * degree + b] - R1[a] @ D).max() <= 1e-15
    assert np.allclose(w, np.repeat(w1 / (degree * 4 * np.pi), degree), rtol=1e-15, atol=0)
    rp, wp = combined_rotations(degree)
    assert np.array_equal(rp, rot) and np.array_equal(wp, w)


def _small_host(seed=7, natom=12, a=15.0):
    rng = np.random.default_rng(seed)
    L = a * U.angstrom
    pos = (rng.random((natom, 3)) - 0.5) * L
    sig = rng.uniform(2.6, 3.4, natom) * U.angstrom
    eps = rng.uniform(20, 60, natom) * U.boltzmann
    return np.eye(3) * L, pos, sig, eps


def _guest(name):
    chk = ChkSystem.from_file(os.path.join(INPUTS, '%s_AA.chk' % name))
    ff = read_pars_file(os.path.join(INPUTS, '%s_pars_vdw_AA_trappe.txt' % name))
    gs = np.array([ff[chk.ffatypes[t]][0] for t in chk.ffatype_ids])
    ge = np.array([ff[chk.ffatypes[t]][1] for t in chk.ffatype_ids])
    return chk, gs, ge


def test_one_site_guest_is_the_spherical_field():
    rvecs, pos, sig, eps = _small_host()
    rng = np.random.default_rng(3)
    pts = (rng.random((40, 3)) - 0.5) * rvecs[0, 0]
    sf, ef = 3.73 * U.angstrom, 148 * U.boltzmann
    beta = 1 / (U.boltzmann * 300.0)
    # the guest keeps its checkpoint-frame centre of mass as an offset: a
    # one-site guest at x0 sees the spherical field at p + x0
    x0 = np.array([[0.3, -0.2, 0.1]])
    ref = O.external_potential(pts + x0, sig, eps, pos, rvecs, sf, ef, cutoff=6 * U.angstrom)
    for degree in (3, 9):
        got = O.effective_potential(pts, x0, [16.0], [sf], [ef], sig, eps, pos, rvecs, beta,
                                    degree=degree, cutoff=6 * U.angstrom, float32=False)
        # the weights sum to 1 up to their 15-decimal rounding
        wsum = O.orientations(degree)[1].sum()
        ok = ref < 1.0
        assert np.abs(got[ok] - (ref[ok] - np.log(wsum) / beta)).max() <= 1e-13 * max(1.0, np.abs(ref[ok]).max())


@pytest.mark.parametrize('name,degree', [('co2', 3), ('c2h6', 5)])
def test_explicit_orientation_loop(name, degree):
    # a plain loop over orientations and atoms with the combined matrix built
    # as R1_a diag(c - s, s + c, 1) and the LJ pair sum written out
    rvecs, pos, sig, eps = _small_host()
    chk, gs, ge = _guest(name)
    rng = np.random.default_rng(11)
    pts = (rng.random((3, 3)) - 0.5) * rvecs[0, 0]
    beta = 1 / (U.boltzmann * 300.0)
    rc = 7 * U.angstrom
    got = O.effective_potential(pts, chk.pos, chk.masses, gs, ge, sig, eps, pos, rvecs, beta, degree=degree,
                                cutoff=rc, float32=False)
    R1, w1 = O.rotation_matrix(degree, 3)
    L = rvecs[0, 0]
    com = (chk.pos * chk.masses[:, None]).sum(0) / chk.masses.sum()
    for ip, p in enumerate(pts):
        terms = []
        for a in range(len(R1)):
            for b in range(degree):
                th = 2 * math.pi * b / degree
                Rab = R1[a] @ np.diag([math.cos(th) - math.sin(th), math.sin(th) + math.cos(th), 1.0])
                U_ = 0.0
                for i in range(len(chk.pos)):
                    q = p + com + Rab @ (chk.pos[i] - com)
                    for j in range(len(pos)):
                        d = q - pos[j]
                        d = d - L * np.round(d / L)
                        R = math.sqrt(float(d @ d) + 1e-16)
                        if R < rc:
                            sm, em = 0.5 * (sig[j] + gs[i]), math.sqrt(eps[j] * ge[i])
                            U_ += 4 * em * ((sm / R) ** 12 - (sm / R) ** 6) - 4 * em * ((sm / rc) ** 12 - (sm / rc) ** 6)
                terms.append((-beta * U_, w1[a] / (degree * 4 * math.pi)))
        mx = max(t for t, _ in terms)
        v = -(mx + math.log(sum(w * math.exp(t - mx) for t, w in terms))) / beta
        assert abs(got[ip] - v) <= 1e-12 * max(abs(v), 1e-3), (got[ip], v)


def test_temperature_limits():
    rvecs, pos, sig, eps = _small_host()
    chk, gs, ge = _guest('co2')
    rng = np.random.default_rng(5)
    pts = (rng.random((6, 3)) - 0.5) * rvecs[0, 0]
    rc = 7 * U.angstrom
    degree = 5
    rot, w = O.orientations(degree)
    com = (chk.pos * chk.masses[:, None]).sum(0) / chk.masses.sum()
    off = np.einsum('rij,nj->rni', rot, chk.pos - com) + com
    Ur = np.zeros((len(pts), len(rot)))
    for i in range(len(chk.pos)):
        q = (pts[:, None, :] + off[None, :, i, :]).reshape(-1, 3)
        Ur += O.external_potential(q, sig, eps, pos, rvecs, gs[i], ge[i], cutoff=rc).reshape(len(pts), -1)
    ok = Ur.max(axis=1) < 0.05
    # high temperature: V -> weighted mean, to first order in beta
    beta = 1e-3 / (U.boltzmann * 300.0)
    got = O.effective_potential(pts, chk.pos, chk.masses, gs, ge, sig, eps, pos, rvecs, beta, degree=degree, cutoff=rc,
                                float32=False)
    mean = (Ur * w).sum(1) / w.sum()
    var = (w * (Ur - mean[:, None]) ** 2).sum(1) / w.sum()
    dev = got - (mean - np.log(w.sum()) / beta)
    assert np.all(np.abs(dev[ok] + 0.5 * beta * var[ok]) <= 0.5 * beta * var[ok] + 1e-14)
    # low temperature: V -> min_r U_r + kT log(1 / w at the minimum), bounded by the weights
    beta = 1e3 / (U.boltzmann * 300.0)
    got = O.effective_potential(pts, chk.pos, chk.masses, gs, ge, sig, eps, pos, rvecs, beta, degree=degree, cutoff=rc,
                                float32=False)
    lo = Ur.min(1) - np.log(w.sum()) / beta
    hi = Ur.min(1) - np.log(w.min()) / beta
    assert np.all(got >= lo - 1e-15) and np.all(got <= hi + 1e-15)


def test_float32_path_agrees_with_float64():
    # the reference's float32 per-atom array against the same sum in float64:
    # the tolerance the float32 path is compared at
    rvecs, pos, sig, eps = _small_host()
    chk, gs, ge = _guest('co2')
    rng = np.random.default_rng(9)
    pts = (rng.random((60, 3)) - 0.5) * rvecs[0, 0]
    beta = 1 / (U.boltzmann * 300.0)
    a = O.effective_potential(pts, chk.pos, chk.masses, gs, ge, sig, eps, pos, rvecs, beta, degree=5,
                              cutoff=7 * U.angstrom, float32=True)
    b = O.effective_potential(pts, chk.pos, chk.masses, gs, ge, sig, eps, pos, rvecs, beta, degree=5,
                              cutoff=7 * U.angstrom, float32=False)
    lim = 1e4 * U.kjmol
    a, b = np.minimum(np.nan_to_num(a, nan=lim, posinf=lim), lim), np.minimum(b, lim)
    assert np.all(np.abs(a - b) <= 4e-7 * np.maximum(np.abs(b), 1 / beta))


def test_dual_model_guest_set_up():
    # system.py:162-229: the multi-site model gives the mass (sum of the
    # atomic masses) and the site count, the spherical model the hard-sphere radius
    from cmmdft_amd.system import DualModelGuest, NonSphericalGuest, SphericalLJGuest
    chk_fn = os.path.join(INPUTS, 'co2_AA.chk')
    par_fn = os.path.join(INPUTS, 'co2_pars_vdw_AA_trappe.txt')
    g = DualModelGuest('CO2', 44.0 * U.amu, 3.73 * U.angstrom, 195.0 * U.boltzmann, chk_fn, par_fn)
    assert isinstance(g, NonSphericalGuest) and isinstance(g, SphericalLJGuest)
    chk = ChkSystem.from_file(chk_fn)
    assert g.natom == 3 and g.mass == float(np.sum(chk.masses))
    assert set(g.ff) == {'O_CD', 'C_CD'}
    g.compute_hardsphere_radius(300.0)
    assert g.Rhs == O.bh_radius(3.73 * U.angstrom, 195.0 * U.boltzmann, 300.0)
    c = g.copy()
    assert c.mass == g.mass and c.sigma == g.sigma and c.natom == 3
    n = NonSphericalGuest('CO2', chk_fn, par_fn)
    with pytest.raises(NotImplementedError):
        n.compute_hardsphere_radius(300.0)


# --- interpolated site potentials (extpot_calculator.py:264-346, 391-548) ---

def test_tricubic_table_matches_reference_checksum():
    import hashlib
    import json
    C = O.tricubic_coefficients()
    assert np.array_equal(C, np.round(C))
    with open(os.path.join(HERE, 'golden', 'tricubic_ref.json')) as fh:
        ref = json.load(fh)
    assert hashlib.sha256(C.astype(np.int8).tobytes()).hexdigest() == ref['sha256_int8']
    assert int(np.count_nonzero(C)) == ref['nonzeros']
    assert [int(v) for v in C.sum(axis=1)] == ref['row_sums']


def _poly_table(rng, n):
    # a random tricubic polynomial in index coordinates and its 8 channels on an n^3 grid
    c = rng.normal(size=(4, 4, 4))
    P = np.polynomial.polynomial
    i = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(i, i, i, indexing='ij')

    def chan(dx, dy, dz):
        cc = c.copy()
        if dx:
            cc = P.polyder(cc, axis=0)
        if dy:
            cc = P.polyder(cc, axis=1)
        if dz:
            cc = P.polyder(cc, axis=2)
        return P.polyval3d(X, Y, Z, cc)
    vals = np.stack([chan(*d) for d in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1),
                                        (0, 1, 1), (1, 1, 1)]])
    return c, vals


def test_tricubic_reproduces_tricubic_polynomials():
    rng = np.random.default_rng(1)
    c, vals = _poly_table(rng, 6)
    h = np.array([0.3, 0.4, 0.5])
    origin = np.array([-1.0, 0.5, 2.0])
    it = O.Interpolator(vals, origin, h)
    s = rng.uniform(0.0, 4.999, (200, 3))           # inside the table, no wrap
    got = it.tricubic(origin + s * h)
    exact = np.polynomial.polynomial.polyval3d(s[:, 0], s[:, 1], s[:, 2], c)
    assert np.abs(got - exact).max() <= 1e-9 * np.abs(exact).max()


def _hermite_eval(vals, origin, h, pos):
    # the device's form: tensor product of cubic Hermite bases, periodic wrap
    n = np.array(vals.shape[1:])
    s = (pos - origin) / h
    f = np.floor(s)
    r = s - f
    i0 = f.astype(int) % n
    i1 = (i0 + 1) % n

    def basis(t):
        return np.array([[2 * t ** 3 - 3 * t ** 2 + 1, t ** 3 - 2 * t ** 2 + t],
                         [-2 * t ** 3 + 3 * t ** 2, t ** 3 - t ** 2]])
    wx, wy, wz = basis(r[:, 0]), basis(r[:, 1]), basis(r[:, 2])
    chans = {(0, 0, 0): 0, (1, 0, 0): 1, (0, 1, 0): 2, (0, 0, 1): 3, (1, 1, 0): 4, (1, 0, 1): 5, (0, 1, 1): 6,
             (1, 1, 1): 7}
    out = np.zeros(len(pos))
    for cx, cy, cz in itertools.product((0, 1), repeat=3):
        ix = (i0, i1)[cx][:, 0]
        iy = (i0, i1)[cy][:, 1]
        iz = (i0, i1)[cz][:, 2]
        for (dx, dy, dz), q in chans.items():
            out += vals[q, ix, iy, iz] * wx[cx, dx] * wy[cy, dy] * wz[cz, dz]
    return out


def test_hermite_form_equals_coefficient_form():
    # the tensor-product Hermite evaluation of the device = the reference's
    # a = X C^T, sum a_ijk x^i y^j z^k, on random channels, with wrap-around
    rng = np.random.default_rng(2)
    vals = rng.normal(size=(8, 5, 6, 7))
    h = np.array([0.25, 0.3, 0.2])
    origin = np.array([0.1, -0.2, 0.3])
    pos = rng.uniform(-3, 3, (300, 3))
    ref = O.Interpolator(vals, origin, h).tricubic(pos)
    got = _hermite_eval(vals, origin, h, pos)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(vals).max()


def test_trilinear_and_estimated_derivatives():
    rng = np.random.default_rng(4)
    n = 24
    i = np.arange(n)
    X, Y, Z = np.meshgrid(i, i, i, indexing='ij')
    k = 2 * np.pi / n
    V = np.cos(k * X) * np.sin(2 * k * Y) + np.cos(k * Z)
    it = O.Interpolator(V, np.zeros(3), np.ones(3))
    est = it.estimate_all_derivatives(V)
    # central differences of a band-limited field: Vx ~ dV/di to O(k^2)
    exact_x = -k * np.sin(k * X) * np.sin(2 * k * Y)
    assert np.abs(est[1] - exact_x).max() <= 0.02 * k
    exact_xyz = np.zeros_like(V)
    assert np.abs(est[7] - exact_xyz).max() <= 1e-12
    # trilinear is exact on nodes and linear in between along an axis
    s = rng.uniform(0, n - 1.001, (50, 3))
    s[:, 1:] = np.floor(s[:, 1:])
    f = np.floor(s[:, 0]).astype(int)
    t = s[:, 0] - f
    lin = (1 - t) * V[f, s[:, 1].astype(int), s[:, 2].astype(int)] + t * V[f + 1, s[:, 1].astype(int),
                                                                            s[:, 2].astype(int)]
    assert np.abs(it.trilinear(s) - lin).max() <= 1e-13
    # 3-D values: tricubic uses the estimated channels
    pos = rng.uniform(-5, 30, (40, 3))
    a = it.tricubic(pos)
    b = O.Interpolator(est, np.zeros(3), np.ones(3)).tricubic(pos)
    assert np.array_equal(a, b)


def test_lj_derivatives_are_derivatives():
    # the channels against central differences of V (and of lower channels)
    rvecs, pos, sig, eps = _small_host()
    rng = np.random.default_rng(6)
    p = (rng.random((30, 3)) - 0.5) * rvecs[0, 0]
    sf, ef = 3.05 * U.angstrom, 79.0 * U.boltzmann
    rc = 7 * U.angstrom
    h = np.array([0.2, 0.3, 0.25])
    base = O.lj_derivatives(p, sig, eps, pos, rvecs, sf, ef, h, cutoff=rc)
    ok = np.abs(base[0]) < 0.01
    d = 1e-4

    def ch(q, shift):
        return O.lj_derivatives(p + shift, sig, eps, pos, rvecs, sf, ef, h, cutoff=rc)[q]
    e = np.eye(3) * d
    for q, (lower, axis) in {1: (0, 0), 2: (0, 1), 3: (0, 2), 4: (1, 1), 5: (1, 2), 6: (2, 2), 7: (4, 2)}.items():
        fd = (ch(lower, e[axis]) - ch(lower, -e[axis])) / (2 * d) * h[axis]
        scale = np.abs(base[q][ok]).max()
        assert np.abs(fd[ok] - base[q][ok]).max() <= 1e-5 * scale, q


def test_coinciding_orientations_are_merged():
    # a linear guest along z: the rotations about z give the same site
    # offsets, so each Lebedev direction is evaluated once with the summed weight
    from cmmdft_amd.extpot import _merge_orientations
    for name, merged in (('co2', True), ('c2h6', False)):
        chk, _, _ = _guest(name)
        for degree in (5, 9):
            rot, w = O.orientations(degree)
            com = (chk.pos * chk.masses[:, None]).sum(0) / chk.masses.sum()
            off = np.einsum('rij,nj->rni', rot, chk.pos - com) + com
            o2, w2 = _merge_orientations(off, w)
            nleb = len(O.lebedev_table(degree)[1])
            assert len(w2) == (nleb if merged else len(w))
            assert abs(w2.sum() - w.sum()) <= 1e-15
            # every original offset set is within rounding of a kept one
            d = np.abs(off[:, None] - o2[None]).max(axis=(2, 3)).min(axis=1)
            assert d.max() <= 1e-12 * np.abs(off).max()
            if merged:
                # the kept weights are the per-direction sums
                assert np.allclose(w2, w.reshape(nleb, degree).sum(axis=1), rtol=1e-15, atol=0)
