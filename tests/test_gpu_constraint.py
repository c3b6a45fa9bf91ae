"""Bulk constraints on the (0,0) mode on the GPU (constraint::treatment
around the implicit solve, suzerain/treatment_constraint.cpp:87-316).

* a small channel: one constrained SMR91 step with the FP64 (zgbsv) solver
  and with the mixed precision (zcgbsvx) solver against the oracle's
  constrained step, per field <= 1e-12; multipliers and the running means
  of the implicit forcing against the oracle's;
* the full 96^3 channel: two constrained steps of the laminar state with
  targets at its own bulk values; the bulks stay on target and the mixed
  precision form agrees with the FP64 form per field <= 1e-12.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SC = {"alpha": 0.5, "beta": 2.0 / 3.0, "gamma": 1.4, "Ma": 1.5, "Pr": 0.7,
      "Re": 300.0}
BULK = {"rho_E": 2.9, "rho_u": 0.25, "rho": 2.45}
PREC = "zcgbsvx,reuse=true,siter=3,diter=5,tolsc=0.1"


def _errors(a, b):
    return [float(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max()) for f in range(5)]


def _small_state(ch, g, ops):
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
    return W


@pytest.mark.parametrize("solver", ["zgbsv", PREC])
def test_small_constrained_step_against_oracle(gpu, solver):
    import torch
    import oracle
    from oracle import channel as och
    from oracle.bspline import breakpoints, integration_coefficients
    from suzerain_amd.channel import Channel, BULK_FIELDS
    print(f"case small constrained step solver={solver}", flush=True)
    ch = Channel(Nx=8, Ny=14, Nz=8, dNx=12, dNz=12, k=6, htdelta=2.0,
                 scenario=SC, solver=solver, bulk=BULK)
    ops = oracle.Ops(bands=ch.operators())
    g = och.Grid(8, 8, 12, 12, ch.Lx, ch.Lz, ops)
    W = _small_state(ch, g, ops)
    weights = integration_coefficients(6, list(breakpoints(14, 6, 0.0, 2.0, 2.0))) / 2.0
    np.testing.assert_allclose(ch.bulk_weights(), weights, rtol=0,
                               atol=1e-13 * weights.max())
    targets = [BULK.get(f, math.nan) for f in BULK_FIELDS]
    ch.set_state(torch.from_numpy(W).to(gpu))
    dt = ch.step(max_dt=2e-3)
    got = ch.get_state(0).cpu().numpy()
    l1y, l2y = ch.lambdas()
    l1x, l1z = math.pi * 8 / ch.Lx, math.pi * 8 / ch.Lz
    scen5 = [SC[q] for q in ("Re", "Pr", "Ma", "alpha", "gamma")]
    L = och.HybridOperator(g, scen5, None, (ch.lower, ch.upper), bulk=targets,
                           weights=weights, Ma=SC["Ma"])
    want, wdt, _ = och.smr91_step(g, W, SC, (ch.lower, ch.upper),
                                  (l1x, l1x * l1x, l1z, l1z * l1z, l1y, l2y),
                                  max_dt=2e-3, L=L)
    errs = _errors(got, want)
    cphi = ch.constraint_multipliers()
    ocphi = L.history[-1][0][:5]
    imp = ch.implicit_means()
    print(f"per-field error {['%.2e' % e for e in errs]}; multipliers "
          f"{cphi} vs {ocphi}", flush=True)
    assert dt == wdt
    assert max(errs) <= 1e-12, errs
    np.testing.assert_allclose(cphi, ocphi, rtol=1e-9, atol=1e-13 * np.abs(ocphi).max())
    for q, name in enumerate(och.IMPLICIT_MEANS):
        ref = L.imp[name]
        tol = 1e-9 * max(np.abs(ref).max(), 1e-300)
        assert np.abs(imp[q] - ref).max() <= tol, (name, imp[q], ref)
    for f, t in enumerate(targets):
        if t == t:
            assert abs(np.dot(weights, got[f, 0, 0].real) - t) <= 1e-13 * abs(t)


def _bulks(W, w):
    return [float(np.dot(w, W[f, 0, 0].real)) for f in range(5)]


def test_full_size_constrained_steps(gpu):
    import torch
    from suzerain_amd.channel import Channel, laminar_state
    cfg = dict(Nx=96, Ny=96, Nz=96, dNx=144, dNz=144, k=8, htdelta=3.0)
    probe = Channel(**cfg, solver="zgbsv")
    W0 = laminar_state(probe)
    w = probe.bulk_weights()
    assert w is None                              # no constraints: no weights
    del probe
    from oracle.bspline import breakpoints, integration_coefficients
    w = integration_coefficients(8, list(breakpoints(96, 8, 0.0, 2.0, 3.0))) / 2.0
    b0 = _bulks(W0, w)
    bulk = {"rho_E": b0[0], "rho_u": b0[1], "rho": b0[4]}
    print(f"case 96^3 constrained, targets {bulk}", flush=True)
    out = {}
    for solver in ("zgbsv", PREC):
        ch = Channel(**cfg, solver=solver, bulk=bulk)
        ch.set_state(torch.from_numpy(W0).to(gpu))
        states = []
        for i in range(2):
            assert ch.step(max_dt=0.02) == 0.02
            s = ch.get_state(0).cpu().numpy()
            states.append(s)
            b = _bulks(s, w)
            dev = [abs(b[f] - b0[f]) / abs(b0[f]) for f in (0, 1, 4)]
            print(f"{solver} step {i}: bulk deviation {dev}, multipliers "
                  f"{ch.constraint_multipliers()}", flush=True)
            assert max(dev) <= 1e-13, dev
        out[solver] = states
        del ch
        torch.cuda.empty_cache()
    for i in range(2):
        errs = _errors(out[PREC][i], out["zgbsv"][i])
        print(f"step {i}: mixed precision vs FP64, per field "
              f"{['%.2e' % e for e in errs]}", flush=True)
        assert max(errs) <= 1e-12, (i, errs)
