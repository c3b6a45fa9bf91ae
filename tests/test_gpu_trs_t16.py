"""Single precision transposed banded solves (CGBTRS 'T' and 'C') at the
band shapes of the channel operators (kl = ku = 34 and 24), where a kernel
with sixteen lanes per system does the work: strided batches, batches with
skipped (singular) systems, and grouped launches mixing system orders in
one wavefront.

The oracle is CGBTRS (oracle/oracle.c) on the same single precision factors.
The kernel sums in a different order, so the bar is a float tolerance on
well conditioned systems (diagonal dominance), 5e-5 relative in the max
norm.  Storage that LAPACK leaves unused (the top-left triangle of the U
rows) is filled with NaN before the solve: the kernel must never use it.
"""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

RTOL = 5e-5
CASES = [(34, 480), (34, 70), (34, 17), (34, 2), (24, 96), (24, 33)]


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _bands(rng, nb, N, kl, ku):
    ldab = 2 * kl + ku + 1
    ab = np.zeros((nb, N, ldab), dtype=np.complex64)
    for s in range(nb):
        for j in range(N):
            lo, hi = max(0, j - ku), min(N, j + kl + 1)
            rows = np.arange(lo, hi)
            ab[s, j, kl + ku + rows - j] = (rng.standard_normal(hi - lo) +
                                            1j * rng.standard_normal(hi - lo))
            ab[s, j, kl + ku] += 3.0 * (kl + ku + 1)
    return ab


def _factor(gpu, ab, kl, ku):
    """GPU CGBTRF, then NaN in the storage LAPACK leaves unused."""
    import suzerain_amd as sa
    abd = _t(ab, gpu)
    ipiv, info = sa.gbtrf(abd, kl, ku, precision="c")
    assert (info.cpu().numpy() == 0).all()
    kv = kl + ku
    N = ab.shape[1]
    for j in range(min(kv, N)):
        abd[:, j, :kv - j] = complex(np.nan, np.nan)
    return abd, ipiv


def _oracle(trans, lu, kl, ku, ipiv, b):
    x = b.copy()
    oracle.zgbtrs(trans, lu.shape[0], kl, ku, lu.reshape(-1).copy(),
                  lu.shape[1], ipiv.copy(), x)
    return x


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("kl,N", CASES)
@pytest.mark.parametrize("trans", ["T", "C"])
def test_strided_matches_oracle(gpu, kl, N, trans):
    import suzerain_amd as sa
    print(f"case kl=ku={kl} N={N} trans={trans}", flush=True)
    rng = np.random.default_rng(100 * kl + N)
    nb = 9
    ab = _bands(rng, nb, N, kl, kl)
    abd, ipiv = _factor(gpu, ab, kl, kl)
    b = (rng.standard_normal((nb, 1, N)) +
         1j * rng.standard_normal((nb, 1, N))).astype(np.complex64)
    bd = _t(b, gpu)
    sa.gbtrs(trans, abd, kl, kl, ipiv, bd, precision="c")
    got = bd.cpu().numpy()
    lu, piv = abd.cpu().numpy(), ipiv.cpu().numpy()
    for s in range(nb):
        want = _oracle(trans, lu[s], kl, kl, piv[s], b[s, 0])
        assert np.isfinite(got[s, 0]).all(), s
        assert _rel(got[s, 0], want) <= RTOL, (s, _rel(got[s, 0], want))


def test_skipped_systems_are_left_alone(gpu):
    import torch
    from suzerain_amd._lib import load
    lib = load()
    f = lib.sz_internal_gbtrs_masked
    vp, i32, i64 = C.c_void_p, C.c_int, C.c_longlong
    f.argtypes = [i32, C.c_char, i32, i32, i32, i32, vp, i32, i64, vp, i64, vp,
                  i32, i64, vp, i32, vp]
    f.restype = i32
    print("case masked kl=ku=34 N=480", flush=True)
    rng = np.random.default_rng(3)
    kl, N, nb = 34, 480, 7
    ab = _bands(rng, nb, N, kl, kl)
    abd, ipiv = _factor(gpu, ab, kl, kl)
    b = (rng.standard_normal((nb, N)) +
         1j * rng.standard_normal((nb, N))).astype(np.complex64)
    bd = _t(b, gpu)
    skip = _t(np.array([0, 1, 0, 0, 1, 1, 0], dtype=np.int32), gpu)
    ldab = abd.shape[2]
    rc = f(1, b"T", N, kl, kl, 1, abd.data_ptr(), ldab, N * ldab,
           ipiv.data_ptr(), N, bd.data_ptr(), N, N, skip.data_ptr(), nb,
           C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    got = bd.cpu().numpy()
    lu, piv, sk = abd.cpu().numpy(), ipiv.cpu().numpy(), skip.cpu().numpy()
    for s in range(nb):
        if sk[s]:
            np.testing.assert_array_equal(got[s], b[s])
        else:
            want = _oracle("T", lu[s], kl, kl, piv[s], b[s])
            assert _rel(got[s], want) <= RTOL, s


@pytest.mark.parametrize("kl,orders", [(34, [480, 70, 2, 17, 400, 35, 96]),
                                       (24, [33, 96, 250, 2, 48])])
def test_grouped_mixed_orders_match_oracle(gpu, kl, orders):
    """Systems of different orders share wavefronts (four per wave), each
    with its own storage, as the ensemble's grouped launches do."""
    import torch
    from suzerain_amd._lib import load
    lib = load()
    f = lib.sz_internal_gbtrs_grouped
    vp, i32 = C.c_void_p, C.c_int
    f.argtypes = [i32, C.c_char, i32, i32, vp, i32, vp, vp, vp, vp, i32, i32, vp]
    f.restype = i32
    print(f"case grouped kl=ku={kl} orders={orders}", flush=True)
    rng = np.random.default_rng(kl + len(orders))
    ldab = 3 * kl + 1
    keep, recs = [], []
    for N in orders:
        ab = _bands(rng, 1, N, kl, kl)
        abd, ipiv = _factor(gpu, ab, kl, kl)
        b = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
        bd = _t(b, gpu)
        keep.append((N, abd, ipiv, b, bd))
        recs.append((N, 0, abd.data_ptr(), ipiv.data_ptr(), bd.data_ptr()))
    gdt = np.dtype([("n", "<i4"), ("pad", "<i4"), ("ab", "<u8"), ("ipiv", "<u8"),
                    ("b", "<u8")])
    gs = np.array(recs, dtype=gdt)
    gsd = _t(gs.view(np.uint8), gpu)
    rc = f(1, b"T", kl, kl, None, ldab, None, None, None, gsd.data_ptr(),
           len(orders), max(orders),
           C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    for s, (N, abd, ipiv, b, bd) in enumerate(keep):
        want = _oracle("T", abd.cpu().numpy()[0], kl, kl, ipiv.cpu().numpy()[0], b)
        got = bd.cpu().numpy()
        assert _rel(got, want) <= RTOL, (s, N, _rel(got, want))


def _split_host(lu, kl, ku):
    """The split layout of one system's LAPACK band storage lu (N, ldab): the
    N columns of U (rows 0..kl+ku) back to back, then the N columns of
    multipliers (rows kl+ku+1..)."""
    U = kl + ku + 1
    return np.concatenate([lu[:, :U].reshape(-1), lu[:, U:].reshape(-1)])


@pytest.mark.parametrize("kl,orders", [(34, [480, 70, 2, 17, 400, 35, 96]),
                                       (24, [33, 96, 250, 2, 48])])
def test_split_factors_convert_and_solve(gpu, kl, orders):
    """Factors converted in place to the split layout by the grouped
    conversion (the listed systems only, bit for bit the host restatement;
    the others untouched), then the grouped solve reading them (descriptor
    layout 1) against CGBTRS on the LAPACK storage."""
    import torch
    from suzerain_amd._lib import load
    lib = load()
    vp, i32 = C.c_void_p, C.c_int
    conv = lib.sz_internal_band_split_grouped
    conv.argtypes = [i32, i32, vp, i32, i32, i32, vp, vp, vp, i32, vp]
    conv.restype = i32
    solve = lib.sz_internal_gbtrs_grouped
    solve.argtypes = [i32, C.c_char, i32, i32, vp, i32, vp, vp, vp, vp, i32, i32, vp]
    solve.restype = i32
    assert lib.sz_internal_gbtrs_split_ok(kl, kl, max(orders)) == 1
    print(f"case split kl=ku={kl} orders={orders}", flush=True)
    rng = np.random.default_rng(7 * kl + len(orders))
    keep, recs = [], []
    for N in orders:
        ab = _bands(rng, 1, N, kl, kl)
        abd, ipiv = _factor(gpu, ab, kl, kl)
        b = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
        lu = abd.cpu().numpy()[0]
        keep.append((N, abd, ipiv, b, _t(b, gpu), lu))
    # two descriptor sets over the same storage: the conversion's (as
    # factored, layout 0) and the solve's (layout 1); the system at
    # position 1 is left out of the list and must stay as it was
    gdt = np.dtype([("n", "<i4"), ("layout", "<i4"), ("ab", "<u8"), ("ipiv", "<u8"),
                    ("b", "<u8")])
    recs = [(N, 1, abd.data_ptr(), ipiv.data_ptr(), bd.data_ptr())
            for N, abd, ipiv, b, bd, lu in keep]
    gs = _t(np.array(recs, dtype=gdt).view(np.uint8), gpu)
    s0 = 100                                  # global numbering of the class
    listed = [s0 + s for s in range(len(orders)) if s != 1]
    lst = _t(np.array(listed[::-1], dtype=np.int32), gpu)
    cnt = _t(np.array([len(listed)], dtype=np.int32), gpu)
    grid = 3
    temp = torch.empty(grid * max(orders) * kl * 2, dtype=torch.float32, device=gpu)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert conv(kl, kl, gs.data_ptr(), s0, len(orders), max(orders), lst.data_ptr(),
                cnt.data_ptr(), temp.data_ptr(), grid, st) == 0
    torch.cuda.synchronize()
    for s, (N, abd, ipiv, b, bd, lu) in enumerate(keep):
        got = abd.cpu().numpy()[0].reshape(-1)
        want = lu.reshape(-1) if s == 1 else _split_host(lu, kl, kl)
        same = (got == want) | (np.isnan(got) & np.isnan(want))
        assert same.all(), (s, N)
    # the unlisted system back in the split layout by the host, then solved
    N1, abd1 = keep[1][0], keep[1][1]
    abd1.copy_(_t(_split_host(keep[1][5], kl, kl).reshape(abd1.shape), gpu))
    assert solve(1, b"T", kl, kl, None, 3 * kl + 1, None, None, None, gs.data_ptr(),
                 len(orders), max(orders), st) == 0
    torch.cuda.synchronize()
    for s, (N, abd, ipiv, b, bd, lu) in enumerate(keep):
        want = _oracle("T", lu, kl, kl, ipiv.cpu().numpy()[0], b)
        got = bd.cpu().numpy()
        assert _rel(got, want) <= RTOL, (s, N, _rel(got, want))
