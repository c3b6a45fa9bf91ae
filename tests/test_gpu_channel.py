"""GPU parity of the explicit operator (pointwise flux, references, step
candidates), the channel's nonlinear operator and one full SMR91 step,
against the oracle (oracle/ns.py, oracle/channel.py; pinned in
tests/test_oracle_ns.py and tests/test_oracle_channel.py).

Floating-point tolerances (FP64), each relative to the largest magnitude of
the compared field, with the measured error printed: pointwise right hand
side and plane means 1e-12; the whole nonlinear operator 1e-12 (the FFTs
differ from numpy's pocketfft only in rounding); one SMR91 step 1e-12, at
8 x 14 x 8 and at the full 96^3 / 144^2 benchmark size; the implicit operator
(assembly, isothermal walls, banded LU, transposed solve) bit-for-bit
against the oracle on the same references, at both sizes.
Size-independent properties are checked at the full size as well: a rest
state stays at rest, a streamwise-homogeneous laminar state keeps every
non-mean mode zero, and the walls satisfy the isothermal constraints after
every step.
"""
import os
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from oracle import channel as och
from oracle import ns as ons

pytestmark = pytest.mark.gpu

SCEN = {"alpha": 0.5, "beta": 2.0 / 3.0, "gamma": 1.4, "Ma": 1.5, "Pr": 0.7,
        "Re": 300.0}


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _scen():
    from suzerain_amd.channel import sz_ns_scenario
    return sz_ns_scenario(*[SCEN[k] for k in ("alpha", "beta", "gamma", "Ma",
                                              "Pr", "Re")])


def _physical_fields(rng, Ny, nz, nx):
    """Random but physical point data: 5 state fields and 41 derivatives."""
    P = np.empty((5, Ny, nz, nx))
    P[4] = rng.uniform(0.8, 1.2, (Ny, nz, nx))
    P[1:4] = rng.uniform(-0.3, 0.3, (3, Ny, nz, nx))
    P[0] = rng.uniform(1.8, 2.6, (Ny, nz, nx))
    A = rng.uniform(-1, 1, (41, Ny, nz, nx))
    return P, A


def test_references_and_rhs_match_oracle(gpu):
    import torch
    import suzerain_amd as sa
    lib = sa.load()
    rng = np.random.default_rng(12)
    Ny, dNz, dNx = 5, 6, 8
    P, A = _physical_fields(rng, Ny, dNz, dNx)
    Pd, Ad = _t(P, gpu), _t(A, gpu)
    Fp = Ny * dNz * dNx
    refs_d = torch.zeros((42, Ny), dtype=torch.float64, device=gpu)
    work = torch.zeros(lib.sz_ns_workspace_bytes(Ny, dNz, dNx) // 8 + 1,
                       dtype=torch.float64, device=gpu)
    sc = _scen()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.sz_ns_references(C.byref(sc), Ny, dNz, dNx, C.c_void_p(Pd.data_ptr()),
                              Fp, C.c_void_p(refs_d.data_ptr()),
                              C.c_void_p(work.data_ptr()), stream)
    assert rc == 0
    refs = och.references(None, SCEN, P)
    got = refs_d.cpu().numpy()
    for q, name in enumerate(ons.REFERENCE_QUANTITIES):
        np.testing.assert_allclose(got[q], refs[name], rtol=1e-13,
                                   atol=1e-14 * np.abs(refs[name]).max(),
                                   err_msg=name)
    Adict = {k: A[i] for i, k in enumerate(och.AUX)}
    l1y = rng.uniform(1, 5, Ny)
    l2y = rng.uniform(1, 25, Ny)
    lxz = np.array([3.0, 4.0, och.SMR91_EVMAXMAG_REAL, och.SMR91_EVMAXMAG_IMAG])
    # device copies held for the whole call (a temporary's memory would be
    # recycled by the allocator as soon as its pointer is taken)
    l1y_d, l2y_d, lxz_d = _t(l1y, gpu), _t(l2y, gpu), _t(lxz, gpu)
    for lin, code in (("rhome_xyz", 1), ("none", 0)):
        out_d = torch.zeros_like(Pd)
        cand_d = torch.zeros(12, dtype=torch.float64, device=gpu)
        rc = lib.sz_ns_rhs(C.byref(sc), code, 1, Ny, dNz, dNx,
                           C.c_void_p(Pd.data_ptr()), Fp,
                           C.c_void_p(Ad.data_ptr()), Fp,
                           C.c_void_p(out_d.data_ptr()), Fp,
                           C.c_void_p(refs_d.data_ptr()),
                           C.c_void_p(l1y_d.data_ptr()),
                           C.c_void_p(l2y_d.data_ptr()),
                           C.c_void_p(lxz_d.data_ptr()),
                           C.c_void_p(cand_d.data_ptr()),
                           C.c_void_p(work.data_ptr()), stream)
        assert rc == 0
        torch.cuda.synchronize()
        lam = (lxz[0], lxz[0] ** 2, lxz[1], lxz[1] ** 2, l1y, l2y)
        want, cand = och.rhs_physical(SCEN, P, Adict, refs, lin, lam,
                                      (lxz[2], lxz[3]))
        out = out_d.cpu().numpy()
        for f in range(5):
            scale = max(np.abs(want[f]).max(), 1e-300)
            assert np.abs(out[f] - want[f]).max() <= 1e-12 * scale, (lin, f)
        np.testing.assert_allclose(cand_d.cpu().numpy(), cand, rtol=1e-13)


@pytest.fixture(scope="module")
def small(gpu):
    from suzerain_amd.channel import Channel
    ch = Channel(Nx=8, Ny=14, Nz=8, dNx=12, dNz=12, k=6, htdelta=2.0,
                 Lx=4 * math.pi, Ly=2.0, Lz=4 * math.pi / 3, scenario=SCEN)
    ops = oracle.Ops(bands=ch.operators())
    g = och.Grid(8, 8, 12, 12, 4 * math.pi, 4 * math.pi / 3, ops)
    return ch, g


def _random_state(g, seed=3, amp=0.004):
    rng = np.random.default_rng(seed)
    Ny = g.Ny
    y = np.linspace(0, 2, Ny)
    M = g.ops.D[0]
    means = [2.5 + 0.2 * (y - 1), 0.4 * y * (2 - y), 0.02 * (y - 1),
             0.1 * y * (2 - y), 1.0 + 0.05 * (y - 1)]
    W = np.zeros((5, g.dNz, g.Xw, Ny), dtype=complex)
    for f in range(5):
        c = (rng.standard_normal((g.dNz, g.Xw, Ny))
             + 1j * rng.standard_normal((g.dNz, g.Xw, Ny))) * amp
        g.zero_dealiasing(c)
        w = g.to_wave(g.to_physical(c)) / (g.dNx * g.dNz)
        g.zero_dealiasing(w)
        w[0, 0] = means[f]
        W[f] = np.linalg.solve(M, w.reshape(-1, Ny).T).T.reshape(w.shape)
    return W


def _lam(ch, g):
    l1y, l2y = ch.lambdas()
    l1x = math.pi * g.Nx / g.Lx
    l1z = math.pi * g.Nz / g.Lz
    return (l1x, l1x * l1x, l1z, l1z * l1z, l1y, l2y)


def test_nonlinear_operator_matches_oracle(gpu, small):
    ch, g = small
    W = _random_state(g)
    ch.set_state(_t(W, gpu), which=1)
    cand = ch.nonlinear(which=1, zeroth=True)
    got = ch.get_state(1).cpu().numpy()
    want, refs, wcand = och.apply_ns(g, W, SCEN, "rhome_xyz", zeroth=True,
                                     lam=_lam(ch, g))
    errs = []
    for f in range(5):
        scale = max(np.abs(want[f]).max(), 1e-300)
        errs.append(np.abs(got[f] - want[f]).max() / scale)
    print(f"nonlinear operator: max rel err per field {errs}")
    assert max(errs) <= 1e-12, errs
    R = ch.refs()
    for q, name in enumerate(ons.REFERENCE_QUANTITIES):
        np.testing.assert_allclose(R[q], refs[name], rtol=1e-12,
                                   atol=1e-13 * np.abs(refs[name]).max(),
                                   err_msg=name)
    np.testing.assert_allclose(cand, wcand, rtol=1e-12)


def _dealiased_modes(ch):
    kx = np.arange(ch.Xw)
    kz = np.arange(ch.dNz)
    wz = np.where(kz < (ch.dNz + 1) // 2, kz, kz - ch.dNz)
    keep = (kx[None, :] <= (ch.Nx - 1) // 2) & (np.abs(wz)[:, None] <= (ch.Nz - 1) // 2)
    return ~keep


def _nonlinear_fused_and_split(ch, W, gpu, monkeypatch):
    out = {}
    for sel in ("fused", "split"):
        if sel == "split":
            monkeypatch.setenv("SZ_DERIV", "split")
        else:
            monkeypatch.delenv("SZ_DERIV", raising=False)
        ch.set_state(_t(W, gpu), which=1)
        cand = ch.nonlinear(which=1, zeroth=True)
        out[sel] = (ch.get_state(1).cpu().numpy(), cand, ch.refs())
    monkeypatch.delenv("SZ_DERIV", raising=False)
    return out


def _fused_vs_split(ch, W, gpu, monkeypatch, seed):
    """The one-launch derivative stage reproduces the per-operator launches
    bit for bit, and both treat dealiased state modes as zero (the
    reference zeroes them before differentiating)."""
    base = _nonlinear_fused_and_split(ch, W, gpu, monkeypatch)
    for k in range(3):
        np.testing.assert_array_equal(base["fused"][k], base["split"][k])
    rng = np.random.default_rng(seed)
    G = W.copy()
    junk = rng.standard_normal(W.shape) + 1j * rng.standard_normal(W.shape)
    G[:, _dealiased_modes(ch), :] = junk[:, _dealiased_modes(ch), :]
    dirty = _nonlinear_fused_and_split(ch, G, gpu, monkeypatch)
    for sel in ("fused", "split"):
        for k in range(3):
            np.testing.assert_array_equal(dirty[sel][k], base["fused"][k])


def test_fused_derivatives_match_split_small(gpu, small, monkeypatch):
    ch, g = small
    _fused_vs_split(ch, _random_state(g, seed=11), gpu, monkeypatch, 12)


def test_fused_derivatives_match_split_tall(gpu, monkeypatch):
    """Ny above one block's width: one pencil per block, strided rows."""
    from suzerain_amd.channel import Channel, laminar_state
    ch = Channel(Nx=4, Ny=300, Nz=6, dNx=6, dNz=9, k=6, htdelta=2.0,
                 scenario=SCEN)
    _fused_vs_split(ch, laminar_state(ch, amp=1e-2, kmax=3), gpu, monkeypatch, 13)


def test_fused_derivatives_match_split_full_size(gpu, big, monkeypatch):
    from suzerain_amd.channel import laminar_state
    _fused_vs_split(big, laminar_state(big, amp=1e-2), gpu, monkeypatch, 14)


def _refs_dict(ch):
    R = ch.refs()
    return {name: R[q] for q, name in enumerate(ons.REFERENCE_QUANTITIES)}


def test_operator_apply_and_invert_round_trip(gpu, small):
    """invert(phi) after apply(phi) recovers the state on every retained
    wavenumber whose walls already satisfy the isothermal constraints.  The
    device results equal the oracle's apply and invert on the same
    references bit for bit; the recovery itself is bounded by the operator's
    conditioning (measured error printed, gate 1e-12)."""
    ch, g = small
    W = _random_state(g, seed=5)
    gamma = SCEN["gamma"]
    Ef = 1.0 / (gamma * (gamma - 1))
    # make the wall coefficients consistent: rho E = Ef rho, m = 0
    for y in (0, g.Ny - 1):
        W[0, :, :, y] = Ef * W[4, :, :, y]
        W[1:4, :, :, y] = 0
    ch.set_state(_t(W, gpu), which=0)
    ch.nonlinear(which=0, zeroth=True)          # collects refs from W
    ch.set_state(_t(W, gpu), which=0)
    phi = 2.5e-3
    ch.apply_operator(phi, which=0)
    applied = ch.get_state(0).cpu().numpy()
    ch.invert_operator(phi, which=0)
    got = ch.get_state(0).cpu().numpy()
    scen5 = [SCEN[k] for k in ("Re", "Pr", "Ma", "alpha", "gamma")]
    R = ons.imexop_refs(_refs_dict(ch))
    bc = ((1.0, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0))
    want_applied = och.apply_operator(g, phi, scen5, R, W)
    np.testing.assert_array_equal(applied[:, g.keep], want_applied[:, g.keep])
    want = och.invert_operator(g, phi, scen5, R, applied, bc)
    np.testing.assert_array_equal(got, want)
    # walls are replaced by the constraint rows, so compare everywhere
    err = np.abs(got - W).max() / np.abs(W).max()
    print(f"apply/invert round trip: max rel err {err:.2e}")
    assert err <= 1e-12


def test_step_matches_oracle(gpu, small):
    ch, g = small
    W = _random_state(g, seed=7)
    ch.set_state(_t(W, gpu), which=0)
    dt = ch.step(max_dt=2e-3)
    got = ch.get_state(0).cpu().numpy()
    bc = ((1.0, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0))
    want, wdt, _ = och.smr91_step(g, W, SCEN, bc, _lam(ch, g), max_dt=2e-3)
    assert dt == wdt == 2e-3
    errs = _field_errors(got, want)
    print(f"SMR91 step 8x14x8: max rel err per field {errs}")
    assert max(errs) <= 1e-12, errs
    # candidates-limited step: same dt selection rule as the oracle
    ch.set_state(_t(W, gpu), which=0)
    dt2 = ch.step(max_dt=0.0)
    _, wdt2, _ = och.smr91_step(g, W, SCEN, bc, _lam(ch, g), max_dt=0.0)
    assert dt2 == pytest.approx(wdt2, rel=1e-11)


def _field_errors(got, want):
    return [float(np.abs(got[f] - want[f]).max() / max(np.abs(want[f]).max(), 1e-300))
            for f in range(5)]


@pytest.fixture(scope="module")
def big(gpu):
    from suzerain_amd.channel import Channel
    return Channel(Nx=96, Ny=96, Nz=96, dNx=144, dNz=144, k=8, htdelta=3.0)


def test_full_size_rest_state_is_steady(gpu, big):
    ch = big
    dt, kl, ku, max_kl, max_ku = ch.operators()
    n = ch.Ny
    Md = np.zeros((n, n))
    for r in range(n):
        for c in range(max(0, r - ku[0]), min(n, r + kl[0] + 1)):
            Md[r, c] = dt[0, r, max_ku + c - r]
    ones = np.linalg.solve(Md, np.ones(n))
    gamma = ch.scenario["gamma"]
    W = np.zeros(ch.shape, dtype=complex)
    W[0, 0, 0] = ones / (gamma * (gamma - 1))
    W[4, 0, 0] = ones
    ch.set_state(_t(W, gpu))
    for _ in range(2):
        ch.step(max_dt=1e-3)
    got = ch.get_state(0).cpu().numpy()
    np.testing.assert_allclose(got, W, rtol=0, atol=1e-12)


def test_full_size_laminar_stays_homogeneous_and_walls_hold(gpu, big):
    from suzerain_amd.channel import laminar_state
    ch = big
    W = laminar_state(ch, amp=0.0)
    ch.set_state(_t(W, gpu))
    dt = ch.step()
    assert 0 < dt < 1
    got = ch.get_state(0).cpu().numpy()
    assert np.isfinite(got).all()
    mean = np.abs(got[:, 0, 0]).max()
    others = got.copy()
    others[:, 0, 0] = 0
    assert np.abs(others).max() <= 1e-13 * mean
    W = laminar_state(ch)                       # perturbed benchmark state
    ch.set_state(_t(W, gpu))
    for _ in range(2):
        ch.step()
    got = ch.get_state(0).cpu().numpy()
    assert np.isfinite(got).all()
    gamma = ch.scenario["gamma"]
    Ef = 1.0 / (gamma * (gamma - 1))
    scale = np.abs(got).max()
    for y in (0, ch.Ny - 1):
        assert np.abs(got[0, :, :, y] - Ef * got[4, :, :, y]).max() < 1e-12 * scale
        assert np.abs(got[1:4, :, :, y]).max() < 1e-12 * scale


def _big_grid(ch):
    ops = oracle.Ops(bands=ch.operators())
    return och.Grid(ch.Nx, ch.Nz, ch.dNx, ch.dNz, ch.Lx, ch.Lz, ops)


def _host_threads():
    return max(1, min(16, os.cpu_count() or 1))


def test_full_size_implicit_solve_matches_oracle(gpu, big):
    """The implicit operator at 96^3 / 144^2: every one of the 4,560 systems
    is assembled, walled, factored and solved on the device; 300 of them,
    spread over the (kx, kz) plane, are compared bit for bit with the oracle
    on the same references."""
    from suzerain_amd.channel import laminar_state
    ch = big
    g = _big_grid(ch)
    W = laminar_state(ch)
    ch.set_state(_t(W, gpu), which=0)
    ch.nonlinear(which=0, zeroth=True)            # references from W
    ch.set_state(_t(W, gpu), which=0)
    phi = -2.5e-3
    ch.invert_operator(phi, which=0)
    got = ch.get_state(0).cpu().numpy()
    systems = list(och._systems(g))
    pick = [systems[i] for i in np.linspace(0, len(systems) - 1, 300).astype(int)]
    scen5 = [ch.scenario[k] for k in ("Re", "Pr", "Ma", "alpha", "gamma")]
    R = ons.imexop_refs(_refs_dict(ch))
    want = och.invert_operator(g, phi, scen5, R, W, (ch.lower, ch.upper),
                               threads=_host_threads(), systems=pick)
    for iz, ix in pick:
        np.testing.assert_array_equal(got[:, iz, ix], want[:, iz, ix],
                                      err_msg=str((iz, ix)))


def test_full_size_step_matches_oracle(gpu, big):
    """One SMR91 step of the benchmark case (96^3, dealiased 144^2, 4,560
    implicit systems per substep) against the oracle's step on host
    threads; the dt choice is identical and every field agrees to 1e-12 of
    its largest magnitude."""
    import time
    from suzerain_amd.channel import laminar_state
    ch = big
    g = _big_grid(ch)
    W = laminar_state(ch)
    ch.set_state(_t(W, gpu), which=0)
    dt = ch.step()
    got = ch.get_state(0).cpu().numpy()
    t0 = time.perf_counter()
    want, wdt, _ = och.smr91_step(g, W, ch.scenario, (ch.lower, ch.upper),
                                  _lam(ch, g), threads=_host_threads())
    errs = _field_errors(got, want)
    print(f"SMR91 step 96^3: dt {dt} (oracle {wdt}), max rel err per field "
          f"{errs}, oracle {time.perf_counter() - t0:.1f} s")
    assert dt == pytest.approx(wdt, rel=1e-12)
    assert max(errs) <= 1e-12, errs
