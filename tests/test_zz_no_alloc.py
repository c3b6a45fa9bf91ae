"""sz_zcgbsvx_batched_noalloc (refinement state in the library's static
scratch) against sz_zcgbsvx_batched (caller workspace) through the C ABI:
every output bit-identical, on two calls in a row with the second batch
smaller than the first, and a batch beyond the scratch refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SZ_ERR_SCRATCH = -1004
NOALLOC_MAX = 131072


def _operator(gpu, nsys):
    import suzerain_amd as sa
    from suzerain_amd.workload import implicit_solve_inputs
    import torch
    inp = implicit_solve_inputs(nsys=nsys, seed=7 + nsys)
    op = sa.BsplineOp.channel(inp["n"], inp["k"], inp["htdelta"])
    S, n = inp["S"], inp["n"]
    kl, ku = max(op.ku[:3]), max(op.kl[:3])
    g = sa.bsmbsm_construct(S, n, kl, ku)
    N, KL, KU, LD = g["N"], g["KL"], g["KU"], g["LD"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    ab = torch.zeros((nsys, N, LD), dtype=torch.complex128, device=gpu)
    sa.model_assemble_ptr(op, S, t(inp["kx"]), t(inp["kz"]), t(inp["cij"]),
                          t(inp["uij"]), inp["phidt"], inp["nu"], ab.data_ptr(),
                          LD, N * LD, nsys, KL, KU)
    b = t(inp["rhs"])
    return ab, b, N, KL, KU, LD


def _solve(gpu, noalloc, ab, b, N, KL, KU, LD, trans, siter, diter, tolsc):
    import torch
    from suzerain_amd._lib import load
    lib = load()
    nsys = ab.shape[0]
    ldafb = 2 * KL + KU + 1
    z = lambda *shape, dt: torch.zeros(shape, dtype=dt, device=gpu)
    out = {
        "fact": torch.full((nsys,), ord("N"), dtype=torch.uint8, device=gpu),
        "apprx": z(nsys, dt=torch.int32),
        "afrob": torch.full((nsys,), -1.0, dtype=torch.float64, device=gpu),
        "afb": z(nsys, N, ldafb, dt=torch.complex128),
        "ipiv": z(nsys, N, dt=torch.int32),
        "x": z(nsys, N, dt=torch.complex128),
        "siter": z(nsys, dt=torch.int32), "diter": z(nsys, dt=torch.int32),
        "tolsc": z(nsys, dt=torch.float64), "res": z(nsys, dt=torch.float64),
        "info": z(nsys, dt=torch.int32), "r": z(nsys, N, dt=torch.complex128),
    }
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [out["fact"].data_ptr(), out["apprx"].data_ptr(), 1, trans.encode(),
            N, KL, KU, ab.data_ptr(), LD, N * LD, out["afrob"].data_ptr(),
            out["afb"].data_ptr(), ldafb, N * ldafb, out["ipiv"].data_ptr(), N,
            b.data_ptr(), N, out["x"].data_ptr(), N, siter, diter, tolsc,
            out["siter"].data_ptr(), out["diter"].data_ptr(),
            out["tolsc"].data_ptr(), out["r"].data_ptr(), N,
            out["res"].data_ptr(), out["info"].data_ptr()]
    if noalloc:
        rc = lib.sz_zcgbsvx_batched_noalloc(*args, nsys, stream)
    else:
        nbytes = lib.sz_zcgbsvx_workspace_bytes(nsys)
        work = torch.zeros((nbytes,), dtype=torch.uint8, device=gpu)
        rc = lib.sz_zcgbsvx_batched(*args, work.data_ptr(), nsys, stream)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("trans,siter,diter,tolsc",
                         [("T", 3, 5, 0.0), ("N", 3, -1, 100.0)])
def test_noalloc_matches_workspace_entry(gpu, trans, siter, diter, tolsc):
    for nsys in (12, 5):            # the second call is the smaller batch
        ab, b, N, KL, KU, LD = _operator(gpu, nsys)
        print(f"no-alloc zcgbsvx: nsys {nsys} N {N} trans {trans} "
              f"siter {siter} diter {diter} tolsc {tolsc}", flush=True)
        want = _solve(gpu, False, ab, b, N, KL, KU, LD, trans, siter, diter,
                      tolsc)
        got = _solve(gpu, True, ab, b, N, KL, KU, LD, trans, siter, diter,
                     tolsc)
        assert (want["info"] == 0).all()
        for k in want:
            if k == "r":
                continue            # a workspace, not an output
            np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_noalloc_refuses_batch_beyond_scratch(gpu):
    import torch
    from suzerain_amd._lib import load
    lib = load()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    nul = None
    rc = lib.sz_zcgbsvx_batched_noalloc(
        nul, nul, 1, b"N", 8, 1, 1, nul, 3, 24, nul, nul, 4, 32, nul, 8, nul, 8,
        nul, 8, 3, 5, 0.0, nul, nul, nul, nul, 8, nul, nul, NOALLOC_MAX + 1,
        stream)
    assert rc == SZ_ERR_SCRATCH
