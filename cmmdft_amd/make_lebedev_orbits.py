"""Writes cmmdft_amd/lebedev_orbits.npz: the Lebedev rules of every degree
the reference tabulates (cmmdft/rotations/lebedev/lebedev_<degree>_<size>.npz,
loaded by AngularGrid, cmmdft/rotations/AngGrid.py:38-71, 256-287), reduced
to their octahedral orbits -- one generator and one weight per orbit, and
the order in which the tables list each orbit type's signed permutations.
``rotations.lebedev`` rebuilds the points and weights from it; this script
checks that the rebuilt rules equal the tables bit for bit before it writes
anything.  Needs the reference checkout; the table travels without it.

Run from the repository root:  python cmmdft_amd/make_lebedev_orbits.py [reference root]
"""
import glob
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
sys.path.insert(0, os.path.dirname(HERE))


def otype(p):
    a = np.abs(p)
    nz = int(np.sum(a == 0.0))
    if nz == 2:
        return 'a1'
    if nz == 1:
        b = np.sort(a)[1:]
        return 'a2' if b[0] == b[1] else 'c'
    s = np.sort(a)
    if s[0] == s[2]:
        return 'a3'
    if s[0] == s[1] or s[1] == s[2]:
        return 'b'
    return 'd'


def signed_perm(g, p):
    """(perm, signs) with p = signs * g[perm] exactly, perm a permutation
    of (0, 1, 2) (the first that fits)."""
    import itertools
    for perm in itertools.permutations(range(3)):
        q = np.abs(g[list(perm)])
        if np.array_equal(q, np.abs(p)):
            s = np.where(p < 0, -1.0, 1.0)
            return list(perm), list(s)
    raise ValueError('no signed permutation maps %r to %r' % (g, p))


def main():
    from cmmdft_amd import rotations as R
    files = glob.glob(os.path.join(REF, 'cmmdft', 'rotations', 'lebedev', 'lebedev_*_*.npz'))
    tables = {}
    for f in files:
        d, n = (int(v) for v in re.findall(r'lebedev_(\d+)_(\d+)\.npz', f)[0])
        with np.load(f, allow_pickle=False) as z:
            tables[(d, n)] = (np.asarray(z['points'], np.float64), np.asarray(z['weights'], np.float64))
    # the rule AngularGrid takes for a degree: the smallest table of it
    rules = {}
    for (d, n) in sorted(tables):
        rules.setdefault(d, (n,) + tables[(d, n)])
    out = {}
    ops = {}
    for d, (n, P, W) in sorted(rules.items()):
        types = [otype(x) for x in P]
        blocks = []
        for i, t in enumerate(types):
            if blocks and blocks[-1][0] == t:
                blocks[-1][2] += 1
            else:
                blocks.append([t, i, 1])
        aw = [np.nan, np.nan, np.nan]
        for t, i0, cnt in blocks:
            if t in ('a1', 'a2', 'a3'):
                aw[int(t[1]) - 1] = W[i0]
                continue
            G = 24 if t in ('b', 'c') else 48
            k = cnt // G
            assert k * G == cnt
            gens = P[i0:i0 + k]
            tab = ops.setdefault(t, [None] * G)
            for j in range(G):
                for o in range(k):
                    perm, s = signed_perm(gens[o], P[i0 + j * k + o])
                    if tab[j] is None:
                        tab[j] = (perm, s)
                    else:
                        q = np.array(tab[j][1]) * gens[o][tab[j][0]]
                        assert np.array_equal(q, P[i0 + j * k + o]), (d, t, j, o)
            # one orbit per row: its first point (the generator) and weight
            w = W[i0:i0 + k * G].reshape(G, k)
            assert np.all(w == w[0]), (d, t)
            out['%s_%d' % (t, d)] = np.column_stack([gens, w[0]])
        out['a_%d' % d] = np.array(aw)
        out['order_%d' % d] = np.array([{'a1': 0, 'a2': 1, 'a3': 2, 'b': 3, 'c': 4, 'd': 5}[b[0]] for b in blocks])
    for t, tab in ops.items():
        out['ops_' + t] = np.array([list(p) + list(s) for p, s in tab], dtype=np.float64)
    out['degrees'] = np.array(sorted(rules))
    out['sizes'] = np.array([rules[d][0] for d in sorted(rules)])
    fn = os.path.join(HERE, 'lebedev_orbits.npz')
    np.savez_compressed(fn, **out)
    # the rebuilt rules equal the tables
    R._ORBITS = None
    for d, (n, P, W) in sorted(rules.items()):
        p, w = R.lebedev(d)
        assert np.array_equal(p, P) and np.array_equal(w, W), d
    print('wrote %s (%d degrees, %d bytes)' % (fn, len(rules), os.path.getsize(fn)))


if __name__ == '__main__':
    main()
