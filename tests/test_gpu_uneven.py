"""Uneven batch: frameworks of different grids solved concurrently with
admission by device footprint, and the batched adjoint (gradients of a
framework's total uptake).  The per-framework batched solver is itself tied
to the plain solver bit for bit and to the oracle (test_gpu_batch)."""
import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W

pytestmark = pytest.mark.gpu

KS = (0, 1, 2)


def _items(npress=4):
    items = []
    for k in KS:
        host = W.host_case5(k)
        # half the resolution of the full workload (0.5 A spacing)
        npts = [W.smooth5(int(round(np.linalg.norm(host.rvecs[a]) / U.angstrom / 0.5))) for a in range(3)]
        prog = W.framework_program(host, npts, images=True)
        items.append((prog, W.chempots_at(W.pressures_uneven(npress))))
    return items


@pytest.fixture(scope='module')
def items():
    return _items()


def test_uneven_batch_equals_separate_batches(items):
    from cmmdft_amd.batch import BatchedAnderson, UnevenBatch
    ref = [BatchedAnderson(p, mu, m=5, damping=0.1, threshold=1e-10, nsteps=400).solve() for p, mu in items]
    ub = UnevenBatch(items, m=5, damping=0.1, threshold=1e-10, nsteps=400)
    res = ub.solve()
    assert sorted(ub.stats['admission_order']) == list(range(len(items)))
    assert ub.stats['peak_concurrent'] == len(items)
    for r, q in zip(res, ref):
        assert np.all(r['converged'])
        assert np.array_equal(r['N'], q['N'])
        assert np.array_equal(r['iters'], q['iters'])


def test_admission_by_footprint(items):
    """A cap that holds only the largest framework: they run one at a time,
    largest first, with the same results."""
    from cmmdft_amd.batch import UnevenBatch
    free = UnevenBatch(items, m=5, damping=0.1, threshold=1e-10, nsteps=400, mem_cap=1 << 40).solve()
    ub = UnevenBatch(items, m=5, damping=0.1, threshold=1e-10, nsteps=400, mem_cap=1)
    with pytest.raises(ValueError):
        ub.solve()
    fps = ub.footprints()
    ub = UnevenBatch(items, m=5, damping=0.1, threshold=1e-10, nsteps=400, mem_cap=max(fps))
    res = ub.solve()
    assert ub.stats['peak_concurrent'] == 1
    assert ub.stats['admission_order'] == sorted(range(len(items)), key=lambda i: -fps[i])
    assert ub.stats['peak_bytes'] <= max(fps)
    for r, q in zip(res, free):
        assert np.array_equal(r['N'], q['N'])


def test_batched_adjoint_matches_single_instance(items):
    import torch
    from cmmdft_amd.adjoint import AdjointSolver, BatchedAdjoint
    from cmmdft_amd.batch import BatchedAnderson
    prog, mus = items[0]
    res = BatchedAnderson(prog, mus, m=5, damping=0.1, threshold=1e-12, nsteps=600).solve()
    assert np.all(res['converged'])
    ba = BatchedAdjoint(prog, tol=1e-11)
    g = ba.gradients(res['rho'])
    single = [AdjointSolver(prog, tol=1e-11).gradients(res['rho'][i]) for i in range(len(mus))]
    for i, s in enumerate(single):
        assert abs(g['dN_dmu'][i] - s['dN_dmu']) <= 1e-9 * abs(s['dN_dmu'])
    tot_s = sum(s['dN_dsigma'] for s in single)
    tot_e = sum(s['dN_depsilon'] for s in single)
    np.testing.assert_allclose(g['dNtot_dsigma'], tot_s, rtol=1e-8)
    np.testing.assert_allclose(g['dNtot_depsilon'], tot_e, rtol=1e-8)
    # ideal-gas-like sanity: uptake grows with mu and with the well depths
    assert np.all(g['dN_dmu'] > 0) and np.all(g['dNtot_depsilon'] > 0)
