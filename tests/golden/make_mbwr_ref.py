"""Writes tests/golden/mbwr_ref.json: the 32 regression coefficients x1..x32
of the reference's modified Benedict-Webb-Rubin equation of state
(cmmdft/eos.py:313-345, Johnson, Zollweg & Gubbins, Mol. Phys. 78, 591
(1993), table 10) as the reference writes them (decimal literals), their
float values and a SHA-256 of the literals joined by newlines.  The oracle's
and the product's tables are checked against it; the mapping of x onto the
a_i(T*), b_i(T*) coefficients is pinned separately by what the published
equation fixes (its critical point, the LJ second virial coefficient).
Needs the reference checkout; reads its source as text only.

Run from the repository root:  python tests/golden/make_mbwr_ref.py [reference root]
"""
import hashlib
import json
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'


def main():
    src = open(os.path.join(REF, 'cmmdft', 'eos.py')).read()
    lit = {}
    for m in re.finditer(r'self\.x(\d+)\s*=\s*([-+0-9.eE]+)', src):
        lit[int(m.group(1))] = m.group(2)
    if sorted(lit) != list(range(1, 33)):
        raise SystemExit('expected the literals x1..x32, found %s' % sorted(lit))
    literals = [lit[i] for i in range(1, 33)]
    out = {'source': 'cmmdft/eos.py:313-345',
           'literals': literals,
           'values': [float(s) for s in literals],
           'sha256_literals': hashlib.sha256('\n'.join(literals).encode()).hexdigest()}
    with open(os.path.join(ROOT, 'tests', 'golden', 'mbwr_ref.json'), 'w') as fh:
        json.dump(out, fh, indent=1)


if __name__ == '__main__':
    main()
