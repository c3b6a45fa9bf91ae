"""Manufactured solutions on the GPU (the nsctpl_rholut isothermal channel of
apps/perfect with alpha = 1, as jobscripts/mms_case.perfect.in runs it).

* sz_mms_evaluate's conserved state and forcing against the oracle's
  symbolic ones at random points, per field <= 1e-12 of the field's size;
* the channel's projection of the exact state (accumulate_manufactured_
  solution) and its L2 norms (compute_field_L2xyz) against the oracle's;
* one forced SMR91 step against the oracle's forced step, per field
  <= 1e-12, and the channel's time advancing by the step;
* the convergence-order gate: one step of 5e-4 from the exact state at
  Ny = 12, 18, 24, 36 (Nx = Nz = 2 Ny / 3, k = 6); the L2 error against the
  exact solution falls at order > 4 in every field, and the mixed
  precision (zcgbsvx) form gives the FP64 (zgbsv) form's errors and
  observed orders (relative 1e-6 and 1e-3).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PREC = "zcgbsvx,reuse=true,siter=3,diter=5,tolsc=0.1"


@pytest.fixture(scope="module")
def oracle_solution():
    from oracle import mms
    P = mms.isothermal_channel()
    P["alpha"] = 1.0
    return P, mms.Solution(P)


def _manufactured():
    from suzerain_amd.mms import Manufactured
    return Manufactured.isothermal_channel().set(alpha=1.0)


def test_pointwise_forcing_matches_oracle(gpu, oracle_solution):
    import torch
    P, sol = oracle_solution
    rng = np.random.default_rng(7)
    n = 4096
    x = rng.uniform(-P["Lx"] / 2, P["Lx"] / 2, n)
    y = rng.uniform(0, P["Ly"], n)
    z = rng.uniform(-P["Lz"] / 2, P["Lz"] / 2, n)
    for t in (0.0, 0.37):
        print(f"case pointwise t={t}", flush=True)
        cons, Q = _manufactured().evaluate(*(torch.from_numpy(a).to(gpu) for a in (x, y, z)), t)
        want_c, want_q = sol.conserved(x, y, z, t), sol.forcing(x, y, z, t)
        for name, got, want in (("conserved", cons, want_c), ("forcing", Q, want_q)):
            got = got.cpu().numpy()
            err = np.abs(got - want).max(axis=1) / np.abs(want).max(axis=1)
            print(f"  {name} per field {err}")
            assert err.max() <= 1e-12, (name, err)


def _channel(N, solver="zgbsv"):
    from suzerain_amd.channel import Channel
    m = _manufactured()
    Nxz = 2 * N // 3
    sc = {q: getattr(m.s, q) for q in ("alpha", "beta", "gamma", "Ma", "Pr", "Re")}
    ch = Channel(Nx=Nxz, Ny=N, Nz=Nxz, dNx=3 * Nxz // 2, dNz=3 * Nxz // 2, k=6,
                 htdelta=0.0, Lx=m.s.Lx, Ly=m.s.Ly, Lz=m.s.Lz, scenario=sc,
                 solver=solver)
    ch.set_manufactured(m)
    return ch


def _oracle_grid(ch):
    import oracle
    from oracle import channel as och
    ops = oracle.Ops(bands=ch.operators())
    return och.Grid(ch.Nx, ch.Nz, ch.dNx, ch.dNz, ch.Lx, ch.Lz, ops)


def _rel(a, b):
    return [float(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max()) for f in range(5)]


def test_projection_and_norms_match_oracle(gpu, oracle_solution):
    from oracle import mms
    from oracle.bspline import breakpoints, gram
    P, sol = oracle_solution
    print("case projection N=18", flush=True)
    ch = _channel(18)
    g = _oracle_grid(ch)
    y = ch.collocation_points()
    ch.manufactured_state(0.25, which=0)
    got = ch.get_state(0).cpu().numpy()
    want = mms.project(g, sol, y, 0.25)
    err = _rel(got, want)
    print(f"projection per field {err}")
    assert max(err) <= 1e-13, err
    G = gram(6, list(breakpoints(18, 6, 0.0, ch.Ly, 0.0)))
    m, f = ch.field_L2(0)
    wm, wf = mms.field_L2(g, want, G)
    print(f"L2 mean {m} vs {wm}; fluct {f} vs {wf}")
    np.testing.assert_allclose(m, wm, rtol=1e-12)
    # fluctuating^2 = total^2 - mean^2 (l2.cpp:141-145) cancels: its bar is
    # relative to total^2
    total2 = wm ** 2 + wf ** 2
    assert (np.abs(f ** 2 - wf ** 2) <= 1e-13 * total2).all(), (f, wf)


def test_forced_step_matches_oracle(gpu, oracle_solution):
    from oracle import channel as och, mms
    P, sol = oracle_solution
    print("case forced step N=12", flush=True)
    ch = _channel(12)
    g = _oracle_grid(ch)
    y = ch.collocation_points()
    ch.set_time(0.1)
    ch.manufactured_state(0.1, which=0)
    W0 = ch.get_state(0).cpu().numpy()
    dt = ch.step(max_dt=5e-4)
    assert dt == 5e-4
    assert ch.time == 0.1 + 5e-4
    got = ch.get_state(0).cpu().numpy()
    l1y, l2y = ch.lambdas()
    l1x, l1z = math.pi * ch.Nx / ch.Lx, math.pi * ch.Nz / ch.Lz
    sc = {q: P[q] for q in ("Re", "Pr", "Ma", "alpha", "beta", "gamma")}
    want, wdt, _ = och.smr91_step(g, W0, sc, (ch.lower, ch.upper),
                                  (l1x, l1x * l1x, l1z, l1z * l1z, l1y, l2y),
                                  max_dt=5e-4, mms=(sol, y), t0=0.1)
    err = _rel(got, want)
    print(f"forced step per field {err}")
    assert wdt == dt
    assert max(err) <= 1e-12, err


def _step_error(N, solver):
    import torch
    ch = _channel(N, solver)
    ch.set_time(0.0)
    ch.manufactured_state(0.0, which=0)
    dt = ch.step(max_dt=5e-4)
    ch.manufactured_state(ch.time, which=1)
    ch.add_scaled(-1.0, 1, 0)                    # state 0 <- state 0 - exact
    m, f = ch.field_L2(0)
    del ch
    torch.cuda.empty_cache()
    return dt, np.hypot(m, f)


def test_convergence_order_unchanged_by_precision_form(gpu):
    Ns = (12, 18, 24, 36)
    errs = {}
    for solver in ("zgbsv", PREC):
        e = []
        for N in Ns:
            dt, err = _step_error(N, solver)
            assert dt == 5e-4
            e.append(err)
            print(f"{solver[:7]} Ny={N}: L2 error per field {err}", flush=True)
        errs[solver] = np.array(e)
    orders = {}
    for solver, e in errs.items():
        orders[solver] = np.array([np.log(e[i] / e[i + 1]) / np.log(Ns[i + 1] / Ns[i])
                                   for i in range(len(Ns) - 1)])
        print(f"{solver[:7]} observed orders\n{orders[solver]}")
    assert orders["zgbsv"].min() > 4, orders["zgbsv"]
    np.testing.assert_allclose(errs[PREC], errs["zgbsv"], rtol=1e-6)
    assert np.abs(orders[PREC] - orders["zgbsv"]).max() <= 1e-3
