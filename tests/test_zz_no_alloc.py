"""The allocation-free LJ gradient entry point (cmm_lj_grid_grad_noalloc)
against cmm_lj_grid_grad with a caller workspace, through the C ABI: two
calls in a row, the second on a smaller grid (the fixed scratch keeps the
first call's partials, which the second must not see), then a size beyond
the scratch, which is rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

from cmmdft_amd import units as U, workloads as W

pytestmark = pytest.mark.gpu


def _inputs(shape, seed):
    import torch
    from cmmdft_amd.extpot import host_atom_table
    host = W.host_case4()
    atoms = torch.from_numpy(np.ascontiguousarray(host_atom_table(host.ff, host))).cuda()
    types = torch.from_numpy(np.ascontiguousarray(host.ffatype_ids.astype(np.int32))).cuda()
    rng = np.random.default_rng(seed)
    w = torch.from_numpy(rng.random((2,) + shape)).cuda()
    return host, atoms, types, w


def _call(lib, name, shape, host, atoms, types, w, work=None):
    import torch
    from cmmdft_amd._lib import _ptr
    rv = np.ascontiguousarray(np.asarray(host.rvecs, dtype=np.float64).reshape(9))
    grad = torch.full((4, 2), np.nan, dtype=torch.float64, device='cuda')
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [shape[0], shape[1], shape[2], rv.ctypes.data_as(C.c_void_p), _ptr(atoms), _ptr(types), atoms.shape[0], 4,
            W.GUEST_SIGMA, W.GUEST_EPSILON, W.CUTOFF, 0, 1e4 * U.kjmol, _ptr(w), 2]
    if work is not None:
        args.append(_ptr(work))
    rc = getattr(lib, name)(*args, _ptr(grad), s)
    torch.cuda.synchronize()
    return rc, grad.cpu().numpy()


def _workspace(lib, shape):
    import torch
    n = C.c_longlong()
    assert lib.cmm_lj_grad_workspace(shape[0], shape[1], shape[2], 4, C.byref(n)) == 0
    return torch.empty(int(n.value), dtype=torch.float64, device='cuda'), int(n.value)


def test_noalloc_gradient_equals_workspace_gradient():
    from cmmdft_amd import _lib
    lib = _lib.load()
    for shape, seed in (((64, 64, 64), 1), ((40, 36, 32), 2)):
        host, atoms, types, w = _inputs(shape, seed)
        work, _ = _workspace(lib, shape)
        rc0, ref = _call(lib, 'cmm_lj_grid_grad', shape, host, atoms, types, w, work)
        rc1, got = _call(lib, 'cmm_lj_grid_grad_noalloc', shape, host, atoms, types, w)
        assert rc0 == 0 and rc1 == 0, lib.cmm_last_error()
        assert np.all(np.isfinite(ref))
        # same kernels and reduction order: identical bits
        assert np.array_equal(got, ref), (got, ref)


def test_noalloc_rejects_grids_beyond_its_scratch():
    from cmmdft_amd import _lib
    lib = _lib.load()
    shape = (256, 256, 256)
    n = C.c_longlong()
    assert lib.cmm_lj_grad_workspace(shape[0], shape[1], shape[2], 8, C.byref(n)) == 0
    assert n.value > (1 << 19)
    # rejected on the host, before any launch (no inputs are touched)
    rc = lib.cmm_lj_grid_grad_noalloc(shape[0], shape[1], shape[2], None, None, None, 0, 8, W.GUEST_SIGMA,
                                      W.GUEST_EPSILON, W.CUTOFF, 0, 0.0, None, 1, None, None)
    assert rc == 1
    assert b'scratch' in lib.cmm_last_error()
