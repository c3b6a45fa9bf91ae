"""GPU parity of the B-spline operator apply and diffwave kernels."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("d", [0, 1, 2, 6])
def test_bsplineop_accumulate_complex(gpu, d):
    import suzerain_amd as sa
    op = sa.BsplineOp.channel(96, 8, 3.0)
    rng = np.random.default_rng(d)
    P = 300
    x = rng.standard_normal((P, op.n)) + 1j * rng.standard_normal((P, op.n))
    y = rng.standard_normal((P, op.n)) + 1j * rng.standard_normal((P, op.n))
    alpha, beta = 0.5 - 2j, 1.5 + 0.25j
    xd, yd = _t(x, gpu), _t(y, gpu)
    op.accumulate(d, alpha, xd, beta, yd)
    got = yd.cpu().numpy()
    want = alpha * (op.dense(d) @ x.T).T + beta * y
    scale = np.abs(want).max()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("dx,dz", [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1),
                                   (0, 2), (3, 1)])
def test_diffwave_accumulate_matches_oracle(gpu, dx, dz):
    import suzerain_amd as sa
    rng = np.random.default_rng(dx * 10 + dz)
    Ny, Nx, Nz = 12, 16, 10
    dNx, dNz = 3 * Nx // 2, 3 * Nz // 2
    X = dNx // 2 + 1
    shape = (dNz, X, Ny)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    y = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    Lx, Lz = 4 * np.pi, 4 * np.pi / 3
    alpha, beta = 1.25 - 0.5j, -0.75 + 2j
    xd, yd = _t(x, gpu), _t(y, gpu)
    sa.diffwave_accumulate(dx, dz, alpha, xd, beta, yd, Lx, Lz, Ny, Nx, dNx,
                           Nz, dNz)
    want = oracle.diffwave_accumulate(dx, dz, alpha, x.copy(), beta, y.copy(),
                                      Lx, Lz, Ny, Nx, dNx, Nz, dNz)
    np.testing.assert_array_equal(yd.cpu().numpy(), want)
    # apply == accumulate with beta = 0
    sa.diffwave_apply(dx, dz, alpha, xd, Lx, Lz, Ny, Nx, dNx, Nz, dNz)
    want0 = oracle.diffwave_accumulate(dx, dz, alpha, x.copy(), 0,
                                       np.zeros_like(x), Lx, Lz, Ny, Nx, dNx,
                                       Nz, dNz)
    np.testing.assert_array_equal(xd.cpu().numpy(), want0)
