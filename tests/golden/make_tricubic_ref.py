"""Writes tests/golden/tricubic_ref.json: a checksum of the reference's 64 x 64
tricubic coefficient table (cmmdft/extpot_calculator.py:20-147, the integer
matrix of Lekien & Marsden 2005 in the reference's corner/channel order),
its number of non-zeros and its row sums -- enough to pin the oracle's
construction of that table (the inverse of the interpolation constraints)
without keeping the table itself.  Needs the reference checkout.

Run from the repository root:  python tests/golden/make_tricubic_ref.py [reference root]
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'


def main():
    src = open(os.path.join(REF, 'cmmdft', 'extpot_calculator.py')).read()
    head = 'coefficients = np.array(['
    i0 = src.index(head) + len(head)
    i1 = src.index('])', i0)
    A = np.array([int(x) for x in re.findall(r'-?\d+', src[i0:i1])], dtype=np.int64).reshape(64, 64)
    out = {'sha256_int8': hashlib.sha256(A.astype(np.int8).tobytes()).hexdigest(),
           'nonzeros': int(np.count_nonzero(A)), 'row_sums': [int(v) for v in A.sum(axis=1)]}
    with open(os.path.join(ROOT, 'tests', 'golden', 'tricubic_ref.json'), 'w') as fh:
        json.dump(out, fh, indent=1)


if __name__ == '__main__':
    main()
