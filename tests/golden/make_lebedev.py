"""Writes tests/golden/lebedev_ref.npz: the points and weights of the
reference's Lebedev tables of degree 3, 5, 7, 9, 11, 13 and 17
(cmmdft/rotations/lebedev/lebedev_<degree>_<size>.npz, the data that
AngularGrid loads, cmmdft/rotations/AngGrid.py:256-287), as fixtures for the
orientational quadrature, and tests/golden/lebedev_digests.json: the SHA-256
of the points and weights (float64 bytes, C order) of the table AngularGrid
takes for every degree it lists (the smallest table of that degree).  Needs
the reference checkout; the fixtures travel without it.

Run from the repository root:  python tests/golden/make_lebedev.py [reference root]
"""
import glob
import hashlib
import json
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
SIZES = {3: 6, 5: 14, 7: 26, 9: 38, 11: 50, 13: 74, 17: 110}


def main():
    out = {}
    for deg, size in SIZES.items():
        fn = os.path.join(REF, 'cmmdft', 'rotations', 'lebedev', 'lebedev_%d_%d.npz' % (deg, size))
        with np.load(fn, allow_pickle=False) as z:
            out['points_%d' % deg] = np.asarray(z['points'], dtype=np.float64)
            out['weights_%d' % deg] = np.asarray(z['weights'], dtype=np.float64)
    np.savez(os.path.join(ROOT, 'tests', 'golden', 'lebedev_ref.npz'), **out)
    tables = {}
    for fn in glob.glob(os.path.join(REF, 'cmmdft', 'rotations', 'lebedev', 'lebedev_*_*.npz')):
        deg, size = (int(v) for v in re.findall(r'lebedev_(\d+)_(\d+)\.npz', fn)[0])
        if deg not in tables or size < tables[deg][0]:
            tables[deg] = (size, fn)
    dig = {}
    for deg, (size, fn) in sorted(tables.items()):
        with np.load(fn, allow_pickle=False) as z:
            p = np.ascontiguousarray(z['points'], dtype=np.float64)
            w = np.ascontiguousarray(z['weights'], dtype=np.float64)
        dig[str(deg)] = {'size': size, 'sha256': hashlib.sha256(p.tobytes() + w.tobytes()).hexdigest()}
    with open(os.path.join(ROOT, 'tests', 'golden', 'lebedev_digests.json'), 'w') as fh:
        json.dump(dig, fh, indent=1)


if __name__ == '__main__':
    main()
