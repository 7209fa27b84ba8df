"""Orientational quadrature for non-spherical guests
(cmmdft/extpot_calculator.py:348-389, cmmdft/rotations/AngGrid.py).

The reference averages the Boltzmann factor of a rigid guest over
orientations: a Lebedev grid on the unit sphere for the direction of the
guest's z axis (``generate_rotation_matrix(degree, 3)``), combined with
``degree`` equally spaced angles about z (``generate_rotation_matrix(degree,
2)``).  Its Lebedev points and weights come from tabulated files
(``rotations/lebedev/lebedev_<degree>_<size>.npz``, degrees 3..131).  Here
the rules of degree 3, 5, 7, 9 and 11 (6, 14, 26, 38 and 50 points) are
generated from their closed-form octahedral orbits (V. I. Lebedev, Zh.
Vychisl. Mat. Mat. Fiz. 16 (1976) 293-306), in the tables' point order, with
the tables' stored values: the weights of the rules of degree >= 7 are
rounded to 15 decimals, and the two irrational orbit coordinates of the 38-
and 50-point rules take the stored doubles (within one ulp of the closed
forms q = sqrt((1 - 1/sqrt(3))/2) and l = 1/sqrt(11)).  The rules of degree
13..131 are rebuilt from their orbits -- one generator and one weight per
orbit, and the order in which the tables list each orbit type's signed
permutations -- kept in ``lebedev_orbits.npz`` (written from the tables by
``make_lebedev_orbits.py``, which checks the rebuilt rules against them bit
for bit).
"""
from __future__ import annotations

import os

import numpy as np

__all__ = ['LEBEDEV_DEGREES', 'lebedev', 'generate_rotation_matrix', 'combined_rotations']

# degree -> number of points of the rules built from closed-form orbits (the
# other degrees the reference lists come from the orbit table)
LEBEDEV_DEGREES = {3: 6, 5: 14, 7: 26, 9: 38, 11: 50}
# every Lebedev degree the reference's tables hold (AngGrid.py:38-71); a
# requested degree maps to the next one up, as AngularGrid does
_ALL_DEGREES = [3, 5, 7, 9, 11, 13, 15, 17, 19, 21, 23, 25, 27, 29, 31, 35, 41, 47, 53, 59, 65, 71, 77, 83, 89,
                95, 101, 107, 113, 119, 125, 131]

_S3 = 0.5773502691896258      # 1/sqrt(3)
_S2 = 0.7071067811865475      # 1/sqrt(2)
_P9 = 0.8880738339771153      # sqrt((1 + 1/sqrt(3))/2)
_Q9 = 0.45970084338098305     # sqrt((1 - 1/sqrt(3))/2), stored value
_L11 = 0.3015113445777635     # 1/sqrt(11), stored value
_M11 = 0.9045340337332909     # 3/sqrt(11)


def _r15(x):
    return float('%.15f' % x)


def _a1():
    return [(1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), (0.0, 0.0, -1.0)]


def _a2():
    a = _S2
    return [(a, a, 0.0), (a, -a, 0.0), (-a, a, 0.0), (-a, -a, 0.0),
            (a, 0.0, a), (a, 0.0, -a), (-a, 0.0, a), (-a, 0.0, -a),
            (0.0, a, a), (0.0, a, -a), (0.0, -a, a), (0.0, -a, -a)]


def _a3():
    a = _S3
    return [(sx * a, sy * a, sz * a) for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)]


def _c1(p, q):
    # (p, q, 0) orbit, in the tables' order
    out = []
    for u, v in ((q, p), (p, q)):
        out += [(u, v, 0.0), (-u, v, 0.0), (-u, -v, 0.0), (u, -v, 0.0)]
    for u, v in ((q, p), (p, q)):
        out += [(u, 0.0, v), (-u, 0.0, v), (-u, 0.0, -v), (u, 0.0, -v)]
    for u, v in ((q, p), (p, q)):
        out += [(0.0, u, v), (0.0, -u, v), (0.0, -u, -v), (0.0, u, -v)]
    return out


def _b1(l, m):
    # (l, l, m) orbit, in the tables' order
    out = []
    for sz in (1, -1):
        out += [(l, l, sz * m), (-l, l, sz * m), (l, -l, sz * m), (-l, -l, sz * m)]
    for sy in (1, -1):
        out += [(l, sy * m, l), (-l, sy * m, l), (l, sy * m, -l), (-l, sy * m, -l)]
    for sx in (1, -1):
        out += [(sx * m, l, l), (sx * m, -l, l), (sx * m, l, -l), (sx * m, -l, -l)]
    return out


def _supported_degree(degree):
    if degree is None or degree < 0 or degree > _ALL_DEGREES[-1]:
        raise ValueError('Lebedev degree should be a positive integer <= %d, got %r' % (_ALL_DEGREES[-1], degree))
    return next(x for x in _ALL_DEGREES if x >= degree)


_ORBITS = None


def _orbits():
    """The orbit table of the rules of degree 13..131 (lebedev_orbits.npz,
    written by make_lebedev_orbits.py from the reference's tables)."""
    global _ORBITS
    if _ORBITS is None:
        fn = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'lebedev_orbits.npz')
        with np.load(fn, allow_pickle=False) as z:
            _ORBITS = {k: z[k] for k in z.files}
    return _ORBITS


def _from_orbits(d):
    """A rule rebuilt from its orbits, in the tables' order: the orbit types
    in the order the table lists them (a1, a2, a3 as _a1/_a2/_a3; b (l, l, m),
    c (p, q, 0) and d (r, s, t) as the signed permutations ops_<type> of each
    orbit's generator, the orbits of a type interleaved, orbit index
    fastest)."""
    O = _orbits()
    pts, wts = [], []
    aw = O['a_%d' % d]
    for t in O['order_%d' % d]:
        t = int(t)
        if t <= 2:
            o = (_a1, _a2, _a3)[t]()
            pts += o
            wts += [aw[t]] * len(o)
            continue
        name = 'bcd'[t - 3]
        gens = O['%s_%d' % (name, d)]
        ops = O['ops_' + name]
        perm = ops[:, :3].astype(int)
        sgn = ops[:, 3:]
        g = gens[:, :3]
        block = np.stack([sgn[j] * g[:, perm[j]] for j in range(len(ops))])   # (ops, orbits, 3)
        pts += [tuple(v) for v in block.reshape(-1, 3)]
        wts += list(np.tile(gens[:, 3], len(ops)))
    return np.array(pts, dtype=np.float64), np.array(wts, dtype=np.float64)


def lebedev(degree):
    """(points (N, 3), weights (N,)) of the Lebedev rule of the given degree
    (or the next one up), weights summing to 1 as stored in the reference's
    tables (AngularGrid multiplies them by 4 pi, AngGrid.py:169).  Degrees
    3..11 are built from their closed-form orbits, 13..131 from the orbit
    table."""
    d = _supported_degree(degree)
    if d not in LEBEDEV_DEGREES:
        return _from_orbits(d)
    if d == 3:
        orbits = [(_a1(), 1.0 / 6.0)]
    elif d == 5:
        orbits = [(_a1(), 1.0 / 15.0), (_a3(), 3.0 / 40.0)]
    elif d == 7:
        orbits = [(_a1(), _r15(1.0 / 21.0)), (_a2(), _r15(4.0 / 105.0)), (_a3(), _r15(27.0 / 840.0))]
    elif d == 9:
        orbits = [(_a1(), _r15(1.0 / 105.0)), (_a3(), _r15(9.0 / 280.0)), (_c1(_P9, _Q9), _r15(1.0 / 35.0))]
    else:
        orbits = [(_a1(), _r15(4.0 / 315.0)), (_a2(), _r15(64.0 / 2835.0)), (_a3(), _r15(27.0 / 1280.0)),
                  (_b1(_L11, _M11), _r15(14641.0 / 725760.0))]
    pts = np.array([p for o, _ in orbits for p in o], dtype=np.float64)
    w = np.array([wt for o, wt in orbits for _ in o], dtype=np.float64)
    return pts, w


def generate_rotation_matrix(degree, dimension):
    """Rotation matrices and weights (extpot_calculator.py:348-389).

    dimension 2: ``degree`` rotations about z by theta_b = 2 pi b / degree,
    weight 1 / (4 pi degree) (a scalar).  dimension 3: one rotation per
    Lebedev point, R = [[c1 c2, -s2, s1 c2], [c1 s2, c2, s1 s2], [-s1, 0, c1]]
    with the point's polar angle phi1 = arctan2(sqrt(x^2 + y^2), z) and
    azimuth phi2 = arctan2(y, x) (so R e_z is the point), weights 4 pi w."""
    if dimension == 2:
        theta = np.linspace(0, 2 * np.pi, degree, endpoint=False)
        c, s = np.cos(theta), np.sin(theta)
        z, o = np.zeros_like(c), np.ones_like(c)
        rot = np.array([[c, -s, z], [s, c, z], [z, z, o]])
        return rot.transpose(2, 0, 1), 1 / (degree * 4 * np.pi)
    if dimension == 3:
        xyz, w = lebedev(degree)
        phi1 = np.arctan2(np.sqrt(xyz[:, 1] ** 2 + xyz[:, 0] ** 2), xyz[:, 2])
        phi2 = np.arctan2(xyz[:, 1], xyz[:, 0])
        c1, s1 = np.cos(phi1), np.sin(phi1)
        c2, s2 = np.cos(phi2), np.sin(phi2)
        z = np.zeros(len(phi1))
        rot = np.array([[c1 * c2, -s2, s1 * c2], [c1 * s2, c2, s1 * s2], [-s1, z, c1]])
        return rot.transpose(2, 0, 1), w * 4 * np.pi
    raise NotImplementedError('rotation grids of dimension %r: only 2 and 3 are used by the effective potential'
                              % (dimension,))


def combined_rotations(degree):
    """The (nrot, 3, 3) orientation matrices and (nrot,) weights of
    ``generate_effective_potential`` (extpot_calculator.py:627-631).

    The reference combines the two sets with ``einsum('aij,bjk->abij')``,
    which does not multiply the matrices: it scales column j of the Lebedev
    rotation R1_a by the j-th row sum of the z rotation R2_b,
    (cos - sin, sin + cos, 1).  That product is kept as it is (it only
    matters for guests not aligned with z), and the weights are
    w1_a w2 repeated over b."""
    R1, w1 = generate_rotation_matrix(degree, 3)
    R2, w2 = generate_rotation_matrix(degree, 2)
    rot = np.einsum('aij,bjk->abij', R1, R2).reshape(-1, 3, 3)
    w = np.repeat(w1 * w2, len(R2))
    return np.ascontiguousarray(rot), w
