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


def test_oracle_reproduces_effpot_fixtures():
    """The oracle's orientational average, LJ derivative table and
    interpolators still give the committed numbers
    (tests/golden/make_golden_effpot.py)."""
    import importlib.util
    import os
    fn = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'make_golden_effpot.py')
    spec = importlib.util.spec_from_file_location('make_golden_effpot', fn)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    new = mod.compute()
    ref = np.load(os.path.join(os.path.dirname(fn), 'oracle_effpot.npz'))
    for k in ref.files:
        a, b = np.asarray(new[k]), ref[k]
        assert a.shape == b.shape, k
        if k.endswith('_f32'):
            assert np.array_equal(a, b), k          # the float32 path is deterministic too
        else:
            scale = max(np.abs(b).max(), 1e-300)
            assert np.abs(a - b).max() <= 1e-12 * scale, k
