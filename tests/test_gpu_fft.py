"""HIP 3-D R2C/C2R transforms against numpy.fft (radix 2/3/4/5/8, odd and
even lengths, batched instances)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(8, 8, 8), (12, 10, 15), (16, 16, 16), (9, 15, 25), (27, 20, 6), (48, 48, 48), (32, 24, 40),
          (64, 72, 75), (81, 50, 36), (128, 128, 128), (160, 96, 100)]


@pytest.mark.parametrize('shape', SHAPES)
def test_r2c_c2r_match_numpy(shape):
    import torch
    from cmmdft_amd._lib import Plan
    rng = np.random.default_rng(7)
    nb = 2
    x = rng.standard_normal((nb,) + shape)
    plan = Plan(shape, np.eye(3) * 20.0, nbatch=nb)
    xd = torch.from_numpy(x).cuda()
    nzh = shape[2] // 2 + 1
    X = torch.empty((nb,) + shape[:2] + (nzh,), dtype=torch.complex128, device='cuda')
    plan.fft_r2c(xd, torch.view_as_real(X))
    ref = np.fft.rfftn(x, axes=(1, 2, 3))
    got = X.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-13 * scale
    back = torch.empty_like(xd)
    plan.fft_c2r(torch.view_as_real(X), back)
    np.testing.assert_allclose(back.cpu().numpy() / np.prod(shape), x, rtol=0, atol=1e-13 * np.abs(x).max())
    # the input spectrum must be untouched by c2r
    assert np.array_equal(X.cpu().numpy(), got)


@pytest.mark.parametrize('shape', [(24, 20, 30), (64, 48, 40), (128, 96, 100)])
def test_every_radix_order_matches_numpy(shape):
    """Each axis through every radix ordering the autotuner may pick."""
    import torch
    from cmmdft_amd._lib import Plan
    from cmmdft_amd.autotune import radix_candidates
    rng = np.random.default_rng(9)
    x = rng.standard_normal(shape)
    plan = Plan(shape, np.eye(3) * 20.0, nbatch=1)
    ref = np.fft.rfftn(x)
    scale = np.abs(ref).max()
    xd = torch.from_numpy(x).cuda()
    nzh = shape[2] // 2 + 1
    X = torch.empty(shape[:2] + (nzh,), dtype=torch.complex128, device='cuda')
    back = torch.empty_like(xd)
    defaults = [plan.radix(a) for a in range(3)]
    for axis in range(3):
        for seq in radix_candidates(shape[axis]):
            plan.set_radix(axis, seq)
            assert plan.radix(axis) == list(seq)
            plan.fft_r2c(xd, torch.view_as_real(X))
            assert np.abs(X.cpu().numpy() - ref).max() <= 1e-13 * scale, (axis, seq)
            plan.fft_c2r(torch.view_as_real(X), back)
            err = np.abs(back.cpu().numpy() / np.prod(shape) - x).max()
            assert err <= 1e-13 * np.abs(x).max(), (axis, seq)
        plan.set_radix(axis, defaults[axis])
    with pytest.raises(ValueError):
        plan.set_radix(0, [2] + plan.radix(0))      # product 2 nx
    with pytest.raises(ValueError):
        plan.set_radix(0, [7, 2])
