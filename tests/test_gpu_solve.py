"""GPU parity of the batched bsmbsm permute/pack and banded LU + solve.

Every case runs the HIP kernels through the C ABI and compares against the
oracle (oracle/, pinned in tests/test_oracle_bsmbsm.py) on identical seeded
inputs: permutations and packs bit-exact, factors and solutions within
1e-12 relative (FP64), identical pivot sequences.
"""
import numpy as np
import pytest

import oracle
import oracle.model as om

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("alpha,beta", [(1, 0), (2 - 1j, 0.5 + 0.25j)])
def test_permute_matches_oracle(gpu, trans, alpha, beta):
    import suzerain_amd as sa
    rng = np.random.default_rng(11)
    S, n, nb = 5, 96, 7
    N = S * n
    x = rng.standard_normal((nb, N)) + 1j * rng.standard_normal((nb, N))
    y = rng.standard_normal((nb, N)) + 1j * rng.standard_normal((nb, N))
    xd, yd = _t(x, gpu), _t(y, gpu)
    sa.aPxpby(trans, S, n, alpha, xd, beta, yd)
    got = yd.cpu().numpy()
    for b in range(nb):
        want = y[b].copy()
        oracle.zaPxpby(trans, S, n, alpha, x[b].copy(), beta, want)
        np.testing.assert_array_equal(got[b], want)


def test_pack_matches_oracle_bitwise(gpu):
    import torch
    import suzerain_amd as sa
    rng = np.random.default_rng(5)
    S, n, kl, ku, nb = 3, 17, 4, 3, 4
    KL, KU = S * (kl + 1) - 1, S * (ku + 1) - 1
    LD = KL + 1 + KU
    N = S * n
    b = rng.standard_normal((nb, n, kl + 1 + ku)) + 1j * rng.standard_normal(
        (nb, n, kl + 1 + ku))
    papt = np.full((nb, N * LD), np.nan + 0j)
    pd = _t(papt, gpu)
    bd = _t(b, gpu)
    for ih in range(S):
        for jh in range(S):
            alpha = complex(ih + 1, -jh) if (ih + jh) % 4 else 0
            sa.zpack(S, n, ih, jh, kl, ku, alpha, bd, KL, KU, pd, LD)
            for s in range(nb):
                oracle.zpack(S, n, ih, jh, kl, ku, alpha, b[s].reshape(-1).copy(),
                             kl + 1 + ku, KL, KU, papt[s], LD)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(pd.cpu().numpy(), papt)


def _random_bands(rng, nb, N, kl, ku, dom=0.0):
    ldab = 2 * kl + ku + 1
    ab = np.zeros((nb, N, ldab), dtype=np.complex128)
    for s in range(nb):
        for j in range(N):
            lo, hi = max(0, j - ku), min(N, j + kl + 1)
            rows = np.arange(lo, hi)
            ab[s, j, kl + ku + rows - j] = (rng.standard_normal(hi - lo) +
                                            1j * rng.standard_normal(hi - lo))
            ab[s, j, kl + ku] += dom
    return ab


@pytest.mark.parametrize("N,kl,ku", [(1, 0, 0), (9, 2, 1), (60, 9, 4),
                                     (130, 5, 20), (480, 34, 34),
                                     (300, 24, 24), (40, 34, 34), (97, 34, 20)])
def test_gbtrf_gbtrs_match_oracle(gpu, N, kl, ku):
    import suzerain_amd as sa
    rng = np.random.default_rng(N * 7 + kl)
    nb = 6
    ab = _random_bands(rng, nb, N, kl, ku)
    ldab = ab.shape[2]
    abd = _t(ab, gpu)
    ipiv, info = sa.gbtrf(abd, kl, ku)
    lu = abd.cpu().numpy()
    ipiv = ipiv.cpu().numpy()
    assert (info.cpu().numpy() == 0).all()
    for s in range(nb):
        ref = ab[s].reshape(-1).copy()
        rinfo, rpiv = oracle.zgbtrf(N, kl, ku, ref, ldab)
        assert rinfo == 0
        np.testing.assert_array_equal(ipiv[s], rpiv)
        ref = ref.reshape(N, ldab)
        # U rows 0..kl+ku of every column and the multipliers below
        # U rows 0..kl+ku and multipliers: identical rounding convention
        np.testing.assert_array_equal(lu[s], ref)
    for trans in ("N", "T", "C"):
        b = rng.standard_normal((nb, 2, N)) + 1j * rng.standard_normal((nb, 2, N))
        bd = _t(b, gpu)
        sa.gbtrs(trans, abd, kl, ku, _t(ipiv, gpu), bd)
        got = bd.cpu().numpy()
        for s in range(nb):
            ref = ab[s].reshape(-1).copy()
            _, rpiv = oracle.zgbtrf(N, kl, ku, ref, ldab)
            for r in range(2):
                want = oracle.zgbtrs(trans, N, kl, ku, ref, ldab, rpiv,
                                     b[s, r].copy())
                np.testing.assert_array_equal(got[s, r], want, err_msg=str((trans, s, r)))


def test_gbtrf_large_batch_matches_small_batch(gpu):
    """1152 systems (between one and two per SIMD on a 256-CU device, the
    batch shape that runs with wave priority over the pivot chain) factor
    bit-identically to the same systems in a 6-system batch, whose factors
    the test above pins to the oracle."""
    import suzerain_amd as sa
    N, kl, ku = 480, 34, 34
    rng = np.random.default_rng(1152)
    small = _random_bands(rng, 6, N, kl, ku)
    big = np.tile(small, (192, 1, 1))
    sd, bd = _t(small, gpu), _t(big, gpu)
    spiv, sinfo = sa.gbtrf(sd, kl, ku)
    bpiv, binfo = sa.gbtrf(bd, kl, ku)
    slu, blu = sd.cpu().numpy(), bd.cpu().numpy()
    spiv, bpiv = spiv.cpu().numpy(), bpiv.cpu().numpy()
    assert (sinfo.cpu().numpy() == 0).all() and (binfo.cpu().numpy() == 0).all()
    for s in range(big.shape[0]):
        np.testing.assert_array_equal(bpiv[s], spiv[s % 6], err_msg=str(s))
        np.testing.assert_array_equal(blu[s], slu[s % 6], err_msg=str(s))


def test_gbtrf_zero_pivot_info_and_gbsv_skip(gpu):
    import suzerain_amd as sa
    N, kl, ku, nb = 12, 2, 2, 3
    ab = np.zeros((nb, N, 2 * kl + ku + 1), dtype=np.complex128)
    ab[:, :, kl + ku] = 1.0
    ab[1, 5, kl + ku] = 0.0          # system 1: zero pivot in column 6
    b = np.ones((nb, 1, N), dtype=np.complex128) * (2 + 1j)
    abd, bd = _t(ab, gpu), _t(b, gpu)
    _, info = sa.gbsv("N", abd, kl, ku, bd)
    assert list(info.cpu().numpy()) == [0, 6, 0]
    got = bd.cpu().numpy()
    np.testing.assert_array_equal(got[1], b[1])      # skipped, untouched
    np.testing.assert_array_equal(got[0], b[0])      # identity solve


@pytest.mark.parametrize("N,kl,ku", [(100, 34, 34), (60, 24, 24)])
def test_gbtrf_zero_pivots_continue_like_lapack(gpu, N, kl, ku):
    """Exactly zero pivot columns set info to the first one and the
    factorization continues without scaling or update (ZGBTF2)."""
    import suzerain_amd as sa
    rng = np.random.default_rng(N + kl)
    nb = 3
    ab = _random_bands(rng, nb, N, kl, ku)
    ldab = ab.shape[2]
    ab[1, 10, :] = 0.0                # column 10 of system 1 entirely zero
    ab[2, N - 1, :] = 0.0             # last column of system 2
    abd = _t(ab, gpu)
    ipiv, info = sa.gbtrf(abd, kl, ku)
    lu, ipiv, info = abd.cpu().numpy(), ipiv.cpu().numpy(), info.cpu().numpy()
    for s in range(nb):
        ref = ab[s].reshape(-1).copy()
        rinfo, rpiv = oracle.zgbtrf(N, kl, ku, ref, ldab)
        assert info[s] == rinfo
        np.testing.assert_array_equal(ipiv[s], rpiv)
        np.testing.assert_array_equal(lu[s], ref.reshape(N, ldab))
    assert info[1] > 0 and info[2] > 0


def test_single_precision_factor_matches_oracle(gpu):
    import torch
    import suzerain_amd as sa
    rng = np.random.default_rng(23)
    N, kl, ku, nb = 200, 14, 14, 4
    ab = _random_bands(rng, nb, N, kl, ku).astype(np.complex64)
    abd = torch.from_numpy(ab).to(gpu)
    ipiv, info = sa.gbtrf(abd, kl, ku, precision="c")
    assert (info.cpu().numpy() == 0).all()
    lu = abd.cpu().numpy()
    for s in range(nb):
        ref = ab[s].reshape(-1).copy()
        _, rpiv = oracle.zgbtrf(N, kl, ku, ref, ab.shape[2])
        np.testing.assert_array_equal(ipiv.cpu().numpy()[s], rpiv)
        np.testing.assert_array_equal(lu[s], ref.reshape(N, -1))


def test_model_operator_assembly_and_solve(gpu):
    """Model operator (batched-solve workload) on a subset of wavenumbers:
    GPU assembly == oracle 25-pack assembly from the same collocation bands;
    GPU permute + gbsv + unpermute == oracle on the GPU-assembled matrix."""
    import torch
    import suzerain_amd as sa
    from suzerain_amd.workload import implicit_solve_inputs
    inp = implicit_solve_inputs(nsys=40)
    op = sa.BsplineOp.channel(inp["n"], inp["k"], inp["htdelta"])
    S, n = inp["S"], inp["n"]
    kl = max(op.ku[:3])     # D's subdiagonals are D_T's superdiagonals
    ku = max(op.kl[:3])
    KL, KU = S * (kl + 1) - 1, S * (ku + 1) - 1
    N, nsys = S * n, len(inp["kx"])
    lu = torch.zeros((nsys, N, 2 * KL + KU + 1), dtype=torch.complex128,
                     device=gpu)
    sa.model_assemble(op, S, _t(inp["kx"], gpu), _t(inp["kz"], gpu),
                      _t(inp["cij"], gpu), _t(inp["uij"], gpu), inp["phidt"],
                      inp["nu"], lu, KL, KU)
    A = lu.cpu().numpy()
    M, D1, D2 = op.dense(0), op.dense(1), op.dense(2)
    for s in (0, 7, nsys - 1):
        ref, _, _ = om.assemble(inp, M, D1, D2, kl, ku, s)
        band = slice(KL, None)
        np.testing.assert_array_equal(A[s][:, band], ref[:, band])
    rhs = _t(inp["rhs"], gpu)
    pb = torch.zeros((nsys, 1, N), dtype=torch.complex128, device=gpu)
    sa.aPxpby("N", S, n, 1, rhs, 0, pb.view(nsys, N))
    ipiv, info = sa.gbsv("N", lu, KL, KU, pb)
    x = torch.zeros_like(rhs)
    sa.aPxpby("T", S, n, 1, pb.view(nsys, N), 0, x)
    assert (info.cpu().numpy() == 0).all()
    X = x.cpu().numpy()
    for s in range(nsys):
        want, rinfo = om.solve_permuted(A[s], KL, KU, S, n, inp["rhs"][s])
        assert rinfo == 0
        np.testing.assert_array_equal(X[s], want, err_msg=str(s))


@pytest.mark.parametrize("N,kl,ku,dom", [(480, 34, 34, 0.0), (480, 34, 34, 6.0),
                                         (300, 24, 24, 0.0), (300, 24, 24, 5.0),
                                         (97, 34, 20, 0.0), (40, 34, 34, 0.0),
                                         (61, 24, 24, 0.0)])
def test_fused_factor_and_transposed_solve_match_oracle(gpu, N, kl, ku, dom):
    """The implicit substep's solve: factorization with the U^T half of
    ZGBTRS 'T' fused in (sz_internal_zgbtrf_ut), then the L^T sweep
    (sz_internal_zgbtrs_lt), bit-identical to ZGBTRF then ZGBTRS 'T'.  Batches
    with and without a dominant diagonal exercise both the interchange-free
    blocks and the fallback after an interchange."""
    import ctypes as C
    import torch
    from suzerain_amd._lib import load
    lib = load()
    i32, i64, vp = C.c_int, C.c_longlong, C.c_void_p
    f_ut = lib.sz_internal_zgbtrf_ut
    f_ut.restype = i32
    f_ut.argtypes = [i32, i32, i32, vp, i32, i64, vp, i64, vp, vp, i64, i32, vp]
    f_lt = lib.sz_internal_zgbtrs_lt
    f_lt.restype = i32
    f_lt.argtypes = [i32, i32, i32, vp, i32, i64, vp, i64, vp, i64, vp, i32,
                     vp]
    rng = np.random.default_rng(N + 3 * kl + ku + int(dom))
    nb = 6
    ab = _random_bands(rng, nb, N, kl, ku, dom)
    ldab = ab.shape[2]
    b = rng.standard_normal((nb, N)) + 1j * rng.standard_normal((nb, N))
    abd, bd = _t(ab, gpu), _t(b, gpu)
    ipiv = torch.zeros((nb, N), dtype=torch.int32, device=gpu)
    info = torch.zeros((nb,), dtype=torch.int32, device=gpu)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert f_ut(N, kl, ku, abd.data_ptr(), ldab, ldab * N, ipiv.data_ptr(), N,
                info.data_ptr(), bd.data_ptr(), N, nb, stream) == 0
    assert f_lt(N, kl, ku, abd.data_ptr(), ldab, ldab * N, ipiv.data_ptr(), N,
                bd.data_ptr(), N, info.data_ptr(), nb, stream) == 0
    torch.cuda.synchronize()
    lu, piv, x = abd.cpu().numpy(), ipiv.cpu().numpy(), bd.cpu().numpy()
    assert (info.cpu().numpy() == 0).all()
    pivoted = 0
    for s in range(nb):
        ref = ab[s].reshape(-1).copy()
        rinfo, rpiv = oracle.zgbtrf(N, kl, ku, ref, ldab)
        assert rinfo == 0
        pivoted += int((np.asarray(rpiv) != np.arange(1, N + 1)).sum())
        np.testing.assert_array_equal(piv[s], rpiv)
        np.testing.assert_array_equal(lu[s], ref.reshape(N, ldab))
        want = oracle.zgbtrs("T", N, kl, ku, ref, ldab, rpiv, b[s].copy())
        np.testing.assert_array_equal(x[s], want, err_msg=str(s))
    if dom == 0.0:
        assert pivoted > 0
