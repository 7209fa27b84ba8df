"""Writes tests/golden/lebedev_ref.npz: the points and weights of the
reference's Lebedev tables of degree 3, 5, 7, 9 and 11
(cmmdft/rotations/lebedev/lebedev_<degree>_<size>.npz, the data that
AngularGrid loads, cmmdft/rotations/AngGrid.py:256-287), as fixtures for the
orientational quadrature.  Needs the reference checkout; the fixture travels
without it.

Run from the repository root:  python tests/golden/make_lebedev.py [reference root]
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
SIZES = {3: 6, 5: 14, 7: 26, 9: 38, 11: 50}


def main():
    out = {}
    for deg, size in SIZES.items():
        fn = os.path.join(REF, 'cmmdft', 'rotations', 'lebedev', 'lebedev_%d_%d.npz' % (deg, size))
        with np.load(fn, allow_pickle=False) as z:
            out['points_%d' % deg] = np.asarray(z['points'], dtype=np.float64)
            out['weights_%d' % deg] = np.asarray(z['weights'], dtype=np.float64)
    np.savez(os.path.join(ROOT, 'tests', 'golden', 'lebedev_ref.npz'), **out)


if __name__ == '__main__':
    main()
