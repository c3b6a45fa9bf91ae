"""The per-system drop-in (include/suzerain_amd_bsmbsm.hpp) from a C++
caller: tests/dropin/dropin_loop.cpp, compiled with g++ against include/
and linked with the library, packs each system's P A P^T through
solver->PAPT with colStride(), supplies, solves, demands, then flushes once.

* zgbsv: every solution equals the oracle's pack + ZGBTRF + ZGBTRS bit for
  bit (same operator values: the caller evaluates the model operator in the
  same order as the oracle's assembly);
* zcgbsvx (the reference's default specification): normwise backward error
  at most 1e-15 for every system."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle.model as om

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "dropin", "dropin_loop")


def _run(inp, spec, tmp):
    fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    S, n, nsys = inp["S"], inp["n"], len(inp["kx"])
    with open(fin, "wb") as f:
        np.array([S, n, inp["k"], nsys], dtype=np.int32).tofile(f)
        np.array([inp["htdelta"], inp["phidt"], inp["nu"]]).tofile(f)
        np.ascontiguousarray(inp["kx"], dtype=np.float64).tofile(f)
        np.ascontiguousarray(inp["kz"], dtype=np.float64).tofile(f)
        np.ascontiguousarray(inp["cij"], dtype=np.float64).tofile(f)
        np.ascontiguousarray(inp["uij"], dtype=np.float64).tofile(f)
        np.ascontiguousarray(inp["rhs"], dtype=np.complex128).tofile(f)
    print(f"dropin_loop {spec}: {nsys} systems", flush=True)
    res = subprocess.run([BIN, fin, fout, spec], capture_output=True, text=True,
                         timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stderr
    with open(fout, "rb") as f:
        info = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        x = np.fromfile(f, dtype=np.complex128).reshape(nsys, S * n)
    return info, x


@pytest.fixture(scope="module")
def model():
    import suzerain_amd as sa
    from suzerain_amd.workload import implicit_solve_inputs
    inp = implicit_solve_inputs(nsys=40)
    op = sa.BsplineOp.channel(inp["n"], inp["k"], inp["htdelta"])
    kl, ku = max(op.ku[:3]), max(op.kl[:3])
    return inp, op, kl, ku


def test_dropin_zgbsv_matches_oracle(gpu, model):
    inp, op, kl, ku = model
    M, D1, D2 = op.dense(0), op.dense(1), op.dense(2)
    with tempfile.TemporaryDirectory() as tmp:
        info, x = _run(inp, "zgbsv", tmp)
    assert info == 0
    for s in range(len(inp["kx"])):
        ab, KL, KU = om.assemble(inp, M, D1, D2, kl, ku, s)
        want, rinfo = om.solve_permuted(ab, KL, KU, inp["S"], inp["n"], inp["rhs"][s])
        assert rinfo == 0
        np.testing.assert_array_equal(x[s], want, err_msg=str(s))


def test_dropin_zcgbsvx_backward_error(gpu, model):
    inp, op, kl, ku = model
    M, D1, D2 = op.dense(0), op.dense(1), op.dense(2)
    with tempfile.TemporaryDirectory() as tmp:
        info, x = _run(inp, "zcgbsvx", tmp)
    assert info == 0
    S, n = inp["S"], inp["n"]
    worst = 0.0
    for s in range(len(inp["kx"])):
        A = np.zeros((S * n, S * n), dtype=complex)
        for ih in range(S):
            for jh in range(S):
                A[ih * n:(ih + 1) * n, jh * n:(jh + 1) * n] = om.submatrix(
                    ih, jh, M, D1, D2, inp["kx"][s], inp["kz"][s], inp["cij"],
                    inp["uij"], inp["phidt"], inp["nu"])
        r = inp["rhs"][s] - A @ x[s]
        be = np.abs(r).max() / (np.abs(A).sum(1).max() * np.abs(x[s]).max())
        worst = max(worst, be)
    print(f"zcgbsvx drop-in: worst normwise backward error {worst:.2e}")
    assert worst <= 1e-15
