"""GPU parity of the compressible implicit operator M + phi L.

The HIP kernels (assembly of P (M + phi L)^T P^T with fused isothermal wall
rows, matrix-free accumulation, right hand side wall treatment) run through
the C ABI and are compared bit for bit with the oracle's restatement
(oracle/imexop.c, pinned in tests/test_oracle_imexop.py) on identical seeded
inputs and identical B-spline operator values.  The full per-wavenumber
implicit solve (assemble, factor, permute, solve with 'T', unpermute) is then
checked end to end against the oracle, also bit for bit.
"""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SCEN = (3000.0, 0.7, 1.5, 0.0, 1.4)      # Re, Pr, Ma, alpha, gamma
PHI = complex(-2.5e-3, 0.0)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def setup(gpu):
    import suzerain_amd as sa
    from suzerain_amd.imexop import Imexop
    op = sa.BsplineOp.channel(96, 8, 3.0)
    dt = np.stack([op.D_T(d) for d in range(3)])
    ops = oracle.Ops(bands=(dt, op.kl, op.ku, op.max_kl, op.max_ku))
    imx = Imexop(op, *SCEN)
    rng = np.random.default_rng(2026)
    refs = rng.uniform(-1, 1, (26, op.n))
    refs[10] = rng.uniform(1e-3, 2e-3, op.n)            # nu > 0
    km = np.array([0.0, 0.5, 3.0, 11.5, 0.5, 23.0])
    kn = np.array([0.0, 0.0, -1.5, 4.5, 70.5, -70.5])
    return {"op": op, "ops": ops, "imx": imx, "refs": refs, "km": km,
            "kn": kn, "refs_d": _t(refs, gpu), "km_d": _t(km, gpu),
            "kn_d": _t(kn, gpu)}


def _geom(S, op):
    KL, KU = S * (op.max_kl + 1) - 1, S * (op.max_ku + 1) - 1
    return KL, KU, KL + 1 + KU


@pytest.mark.parametrize("positions", [(0, 1, 2, 3, 4), (4, 0, 1, 2, 3),
                                       (0, 1, 2, 3, -1)])
@pytest.mark.parametrize("packf", [False, True])
def test_pack_bitwise(gpu, setup, positions, packf):
    import torch
    op, ops, imx = setup["op"], setup["ops"], setup["imx"]
    S = sum(p >= 0 for p in positions)
    KL, KU, LD = _geom(S, op)
    N = S * op.n
    ld = LD + KL if packf else LD
    nsys = len(setup["km"])
    buf = torch.full((nsys, N * ld), float("nan"), dtype=torch.complex128,
                     device=gpu)
    g = imx.pack(PHI, setup["km_d"], setup["kn_d"], setup["refs_d"],
                 buf.data_ptr(), ld, N * ld, nsys, positions=positions,
                 packf=packf)
    assert (g["KL"], g["KU"], g["S"]) == (KL, KU, S)
    got = buf.cpu().numpy()
    for s in range(nsys):
        want = np.full(N * ld, np.nan + 0j)
        oracle.imexop_pack(PHI, setup["km"][s], setup["kn"][s], SCEN,
                           setup["refs"], ops, want, ld, idx=positions,
                           offset=KL if packf else 0)
        np.testing.assert_array_equal(got[s], want, err_msg=f"system {s}")


def test_pack_with_isothermal_walls_bitwise(gpu, setup):
    import torch
    from suzerain_amd.imexop import isothermal_bc
    op, ops, imx = setup["op"], setup["ops"], setup["imx"]
    S = 5
    KL, KU, LD = _geom(S, op)
    N = S * op.n
    ld = LD + KL
    nsys = len(setup["km"])
    lower, upper = (1.0, 0.0, 0.0, 0.0), (1.25, 0.5, 0.0, -0.25)
    buf = torch.zeros((nsys, N * ld), dtype=torch.complex128, device=gpu)
    imx.pack(PHI, setup["km_d"], setup["kn_d"], setup["refs_d"], buf.data_ptr(),
             ld, N * ld, nsys, bc=isothermal_bc(lower, upper), packf=True)
    got = buf.cpu().numpy()
    for s in range(nsys):
        want = np.zeros(N * ld, dtype=complex)
        oracle.imexop_pack(PHI, setup["km"][s], setup["kn"][s], SCEN,
                           setup["refs"], ops, want, ld, offset=KL)
        oracle.isothermal_op(S, op.n, KL, KU, want, ld, SCEN, lower, upper,
                             offset=KL)
        np.testing.assert_array_equal(got[s], want, err_msg=f"system {s}")


@pytest.mark.parametrize("beta", [0.0, complex(0.5, -0.25)])
def test_accumulate_bitwise(gpu, setup, beta):
    op, ops, imx = setup["op"], setup["ops"], setup["imx"]
    n, nsys = op.n, len(setup["km"])
    rng = np.random.default_rng(7)
    x = rng.standard_normal((5, nsys, n)) + 1j * rng.standard_normal((5, nsys, n))
    y = rng.standard_normal((5, nsys, n)) + 1j * rng.standard_normal((5, nsys, n))
    xd, yd = _t(x, gpu), _t(y, gpu)
    imx.accumulate(PHI, setup["km_d"], setup["kn_d"], setup["refs_d"], xd, beta,
                   yd)
    got = yd.cpu().numpy()
    for s in range(nsys):
        want = np.ascontiguousarray(y[:, s, :])
        oracle.imexop_accumulate(PHI, setup["km"][s], setup["kn"][s], SCEN,
                                 setup["refs"], ops, x[:, s, :], beta, want)
        np.testing.assert_array_equal(got[:, s, :], want, err_msg=f"system {s}")


def test_accumulate_field_major_offsets(gpu, setup):
    """Pencils scattered through per-field arrays with per-system offsets,
    the wave-space state layout; fields may be omitted."""
    import torch
    op, ops, imx = setup["op"], setup["ops"], setup["imx"]
    n, nsys = op.n, len(setup["km"])
    rng = np.random.default_rng(8)
    npencil = 3 * nsys
    offs = np.array(rng.permutation(npencil)[:nsys] * n, dtype=np.int64)
    X = rng.standard_normal((5, npencil * n)) + 1j * rng.standard_normal(
        (5, npencil * n))
    Xd = _t(X, gpu)
    Yd = torch.zeros_like(Xd)
    present = (True, True, True, True, False)
    xs = [Xd[f] if present[f] else None for f in range(5)]
    ys = [Yd[f] if present[f] else None for f in range(5)]
    offd = _t(offs, gpu)
    imx.accumulate(PHI, setup["km_d"], setup["kn_d"], setup["refs_d"], xs, 0.0,
                   ys, nsys=nsys, inoff=offd, outoff=offd)
    got = Yd.cpu().numpy()
    for s in range(nsys):
        o = offs[s]
        xin = np.ascontiguousarray(X[:, o:o + n])
        want = np.zeros((5, n), dtype=complex)
        oracle.imexop_accumulate(PHI, setup["km"][s], setup["kn"][s], SCEN,
                                 setup["refs"], ops, xin, 0.0, want,
                                 present=present)
        np.testing.assert_array_equal(got[:4, o:o + n], want[:4])
        assert not got[4].any()


def test_implicit_solve_end_to_end_bitwise(gpu, setup):
    """invert_mass_plus_scaled_operator for rhome_xyz
    (operator_hybrid_isothermal.cpp:612-689): assemble (packf) + walls, P b
    + wall rhs, ZGBTRF, ZGBTRS('T'), P^T, against the oracle chain."""
    import torch
    import suzerain_amd as sa
    from suzerain_amd.imexop import isothermal_bc, Imexop
    op, ops, imx = setup["op"], setup["ops"], setup["imx"]
    S, n = 5, op.n
    KL, KU, LD = _geom(S, op)
    N = S * n
    ldlu = LD + KL
    nsys = len(setup["km"])
    rng = np.random.default_rng(9)
    b = rng.standard_normal((nsys, N)) + 1j * rng.standard_normal((nsys, N))
    lower, upper = (1.0, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0)
    lu = torch.zeros((nsys, N, ldlu), dtype=torch.complex128, device=gpu)
    imx.pack(PHI, setup["km_d"], setup["kn_d"], setup["refs_d"], lu.data_ptr(),
             ldlu, N * ldlu, nsys, bc=isothermal_bc(lower, upper), packf=True)
    bd = _t(b, gpu)
    pb = torch.zeros((nsys, 1, N), dtype=torch.complex128, device=gpu)
    sa.aPxpby("N", S, n, 1, bd, 0, pb.view(nsys, N))
    Imexop.isothermal_rhs(S, n, pb.data_ptr(), N, nsys)
    ipiv, info = sa.gbtrf(lu, KL, KU)
    sa.gbtrs("T", lu, KL, KU, ipiv, pb)
    x = torch.zeros((nsys, N), dtype=torch.complex128, device=gpu)
    sa.aPxpby("T", S, n, 1, pb.view(nsys, N), 0, x)
    assert int(info.abs().sum()) == 0
    got = x.cpu().numpy()
    for s in range(nsys):
        ab = np.zeros(N * ldlu, dtype=complex)
        oracle.imexop_pack(PHI, setup["km"][s], setup["kn"][s], SCEN,
                           setup["refs"], ops, ab, ldlu, offset=KL)
        oracle.isothermal_op(S, n, KL, KU, ab, ldlu, SCEN, lower, upper,
                             offset=KL)
        p = np.zeros(N, dtype=complex)
        oracle.zaPxpby("N", S, n, 1, b[s].copy(), 0, p)
        oracle.isothermal_rhs(S, n, p)
        inf, piv = oracle.zgbtrf(N, KL, KU, ab, ldlu)
        assert inf == 0
        oracle.zgbtrs("T", N, KL, KU, ab, ldlu, piv, p)
        want = np.zeros(N, dtype=complex)
        oracle.zaPxpby("T", S, n, 1, p, 0, want)
        np.testing.assert_array_equal(got[s], want, err_msg=f"system {s}")
