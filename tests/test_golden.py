"""The oracle against its committed fixtures (tests/golden/oracle_small.npz,
written by tests/golden/make_golden.py): guards the restatement against
drift.  The GPU path is compared with the same fixtures in
test_gpu_golden.py."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))


def test_oracle_reproduces_fixtures():
    import make_golden
    ref = np.load(os.path.join(HERE, 'golden', 'oracle_small.npz'))
    out = make_golden.compute()
    assert set(out) == set(ref.files)
    for k in ref.files:
        a, b = np.asarray(out[k]), ref[k]
        assert a.shape == b.shape, k
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-12 * scale, k
