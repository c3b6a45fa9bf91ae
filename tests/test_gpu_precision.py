"""The channel's mixed precision implicit solve (zcgbsvx specifications) on
the GPU: float factors refined with double residuals of the matrix-free
operator, factors kept per SMR91 substep and reused across steps.

* one step of a small channel against the oracle's FP64 step, <= 1e-12;
* several full-size (96^3) steps against the plain FP64 (zgbsv) path, per
  field <= 1e-12, with the factors of step 1 reused by later steps;
* siter = -1 (double precision only) sends every system to the double
  fallback, which must reproduce the zgbsv path bit for bit;
* a jump in the time step makes the reused factors stall: the systems are
  refactored and restarted (counted) and the result still meets 1e-12;
* the precision form at 192^3 (18,336 systems of order N = 960): two steps
  against the plain FP64 path, per field <= 1e-12, every system within the
  refinement's residual target after at most three sweeps, none sent to
  double precision.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPEC = "zcgbsvx,reuse=true,siter=3,diter=5,tolsc=0.1"


def _small(spec):
    import torch
    from suzerain_amd.channel import Channel
    sc = {"alpha": 0.5, "beta": 2.0 / 3.0, "gamma": 1.4, "Ma": 1.5, "Pr": 0.7,
          "Re": 300.0}
    return Channel(Nx=8, Ny=14, Nz=8, dNx=12, dNz=12, k=6, htdelta=2.0,
                   scenario=sc, solver=spec), sc


def _field_errors(a, b):
    return [float(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max()) for f in range(5)]


def test_small_step_against_oracle(gpu):
    import torch
    import oracle
    from oracle import channel as och
    ch, sc = _small(SPEC)
    ops = oracle.Ops(bands=ch.operators())
    g = och.Grid(8, 8, 12, 12, ch.Lx, ch.Lz, ops)
    rng = np.random.default_rng(5)
    y = np.linspace(0, 2, g.Ny)
    means = [2.5 + 0.2 * (y - 1), 0.4 * y * (2 - y), 0.02 * (y - 1),
             0.1 * y * (2 - y), 1.0 + 0.05 * (y - 1)]
    W = np.zeros((5, g.dNz, g.Xw, g.Ny), dtype=complex)
    for f in range(5):
        c = (rng.standard_normal(W.shape[1:])
             + 1j * rng.standard_normal(W.shape[1:])) * 0.004
        g.zero_dealiasing(c)
        w = g.to_wave(g.to_physical(c)) / (g.dNx * g.dNz)
        g.zero_dealiasing(w)
        w[0, 0] = means[f]
        W[f] = np.linalg.solve(ops.D[0], w.reshape(-1, g.Ny).T).T.reshape(w.shape)
    ch.set_state(torch.from_numpy(W).to(gpu))
    dt = ch.step(max_dt=2e-3)
    got = ch.get_state(0).cpu().numpy()
    l1y, l2y = ch.lambdas()
    l1x, l1z = math.pi * 8 / ch.Lx, math.pi * 8 / ch.Lz
    want, wdt, _ = och.smr91_step(g, W, sc, (ch.lower, ch.upper),
                                  (l1x, l1x * l1x, l1z, l1z * l1z, l1y, l2y),
                                  max_dt=2e-3)
    errs = _field_errors(got, want)
    print(f"mixed precision small step vs oracle: {errs}")
    assert dt == wdt
    assert max(errs) <= 1e-12, errs


def _full(spec):
    from suzerain_amd.channel import Channel
    return Channel(Nx=96, Ny=96, Nz=96, dNx=144, dNz=144, k=8, htdelta=3.0,
                   solver=spec)


@pytest.fixture(scope="module")
def fp64_steps(gpu):
    """States after each of 3 steps (fixed dt 0.02) of the plain FP64 path
    at 96^3, and the initial state."""
    import torch
    from suzerain_amd.channel import laminar_state
    ch = _full("zgbsv")
    W0 = laminar_state(ch)
    ch.set_state(torch.from_numpy(W0).to(gpu))
    out = []
    for _ in range(3):
        ch.step(max_dt=0.02)
        out.append(ch.get_state(0).cpu().numpy())
    del ch
    torch.cuda.empty_cache()
    return W0, out


def test_full_size_steps_match_fp64(gpu, fp64_steps):
    import torch
    W0, want = fp64_steps
    ch = _full(SPEC)
    assert ch.solver_slots() == 3
    ch.set_state(torch.from_numpy(W0).to(gpu))
    for i in range(3):
        ch.step(max_dt=0.02)
        got = ch.get_state(0).cpu().numpy()
        errs = _field_errors(got, want[i])
        si = ch.solver_stats("siter")
        counts = ch.solver_counts()
        print(f"step {i}: per-field error {['%.2e' % e for e in errs]}, "
              f"siter max {int(si.max())}, (factored, restarted, fallback) "
              f"of the last substep {counts}", flush=True)
        assert max(errs) <= 1e-12, (i, errs)
        assert int(si.max()) <= 3
        if i == 0:
            assert counts[0] == ch.nsys           # every system factored
        else:
            assert counts[0] < ch.nsys // 10      # factors reused


def test_double_only_equals_zgbsv(gpu, fp64_steps):
    """siter = -1: no single precision sweeps, every system with a nonzero
    right hand side is solved by the double precision fallback -- the zgbsv
    kernels on the same operator (a Langou tolerance, tolsc > 0, makes only
    an exactly zero right hand side count as solved by x = 0; with tolsc = 0
    the reference also accepts x = 0 when ||b||_2 <= eps)."""
    import torch
    W0, want = fp64_steps
    ch = _full("zcgbsvx,reuse=true,siter=-1,diter=5,tolsc=1")
    ch.set_state(torch.from_numpy(W0).to(gpu))
    ch.step(max_dt=0.02)
    got = ch.get_state(0).cpu().numpy()
    fallback = ch.solver_counts()[2]
    print(f"double fallback: {fallback} of {ch.nsys} systems")
    assert fallback >= ch.nsys // 2
    np.testing.assert_array_equal(got, want[0])


def test_stalled_reuse_restarts(gpu, fp64_steps):
    """Factors made at dt = 0.02, the next step at dt = 0.005: refinement
    with the stale factors stalls, the systems are refactored."""
    import torch
    from suzerain_amd.channel import Channel
    W0, _ = fp64_steps
    ch = _full(SPEC)
    ref = _full("zgbsv")
    for c in (ch, ref):
        c.set_state(torch.from_numpy(W0).to(gpu))
        c.step(max_dt=0.02)
    ref.step(max_dt=0.005)
    ch.step(max_dt=0.005)
    restarted = ch.solver_counts()[1]
    got, want = ch.get_state(0).cpu().numpy(), ref.get_state(0).cpu().numpy()
    errs = _field_errors(got, want)
    print(f"dt jump: restarted {restarted} of {ch.nsys} in the last substep, "
          f"errors {errs}")
    assert restarted > 0
    assert max(errs) <= 1e-12, errs


def test_192_cube_steps_match_fp64(gpu):
    """The precision form at 192^3 (dealiased 288 x 288, k = 8): float
    factors per substep slot, FP64 residual sweeps; the plain FP64 path
    (pinned to the oracle at 96^3 and on small grids) is the reference."""
    import torch
    from suzerain_amd.channel import Channel, laminar_state
    cfg = dict(Nx=192, Ny=192, Nz=192, dNx=288, dNz=288, k=8, htdelta=3.0)
    ref = Channel(**cfg, solver="zgbsv")
    W0 = torch.from_numpy(laminar_state(ref)).to(gpu)
    ref.set_state(W0)
    want = []
    for _ in range(2):
        assert ref.step(max_dt=0.01) == 0.01
        want.append(ref.get_state(0).cpu().numpy())
    nsys = ref.nsys
    del ref
    torch.cuda.empty_cache()
    ch = Channel(**cfg, solver=SPEC)
    assert ch.nsys == nsys == 18336
    ch.set_state(W0)
    for i in range(2):
        assert ch.step(max_dt=0.01) == 0.01
        errs = _field_errors(ch.get_state(0).cpu().numpy(), want[i])
        si = ch.solver_stats("siter")
        ratio = ch.solver_stats("tolsc")          # final residual / its target
        counts = ch.solver_counts()
        print(f"192^3 step {i}: per-field error {['%.2e' % e for e in errs]}, "
              f"siter max {int(si.max())}, residual/target max {ratio.max():.3f}, "
              f"(factored, restarted, fallback) {counts}", flush=True)
        assert max(errs) <= 1e-12, (i, errs)
        assert int(si.max()) <= 3
        assert float(ratio.max()) <= 1.0
        assert counts[2] == 0
    del ch
    torch.cuda.empty_cache()
