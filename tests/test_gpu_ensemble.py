"""Ensembles on the GPU: every member's state after steps in the ensemble
equals the same member stepped alone (a Channel with the same mixed
precision specification and time step) bit for bit, eagerly and replayed
from a captured hipGraph, and the autotuner's trial steps leave the states
as they were."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPEC = "zcgbsvx,reuse=true,siter=3,diter=5,tolsc=0.1"


def _members():
    from suzerain_amd.channel import DEFAULT_SCENARIO
    sc = lambda re, ma: dict(DEFAULT_SCENARIO, Re=re, Ma=ma)
    return [dict(Nx=8, Ny=18, Nz=12, k=8, htdelta=3.0, scenario=sc(1500.0, 0.6)),
            dict(Nx=12, Ny=14, Nz=8, k=6, htdelta=3.0, scenario=sc(4000.0, 1.2)),
            dict(Nx=8, Ny=24, Nz=8, k=8, htdelta=3.0, scenario=sc(2500.0, 0.9)),
            dict(Nx=12, Ny=16, Nz=12, k=6, htdelta=3.0, scenario=sc(5500.0, 0.4))]


def _initial(cfg, gpu):
    import torch
    from suzerain_amd.channel import Channel, laminar_state
    ch = Channel(**cfg, solver="zgbsv")
    W = laminar_state(ch, seed=7 + cfg["Ny"])
    return torch.from_numpy(W).to(gpu)


def _alone(cfg, W, dt, steps, gpu):
    from suzerain_amd.channel import Channel
    ch = Channel(**cfg, solver=SPEC)
    ch.set_state(W)
    for _ in range(steps):
        got = ch.step(max_dt=dt)
        assert got == dt
    return ch.get_state(0).cpu().numpy()


@pytest.mark.parametrize("graph", [False, True])
def test_members_equal_alone(gpu, graph):
    from suzerain_amd.ensemble import Ensemble
    members = _members()
    W = [_initial(c, gpu) for c in members]
    dts = [2e-3, 1.5e-3, 2.5e-3, 1e-3]
    ens = Ensemble(members, dt=dts, solver=SPEC, graph=graph)
    for m in range(len(members)):
        ens.set_state(m, W[m])
    steps = 3
    ens.step(steps)
    ens.check()
    for m, c in enumerate(members):
        cand = ens.member(m).candidates()
        assert dts[m] <= np.nanmin(cand), (m, dts[m], np.nanmin(cand))
        want = _alone(c, W[m], dts[m], steps, gpu)
        got = ens.get_state(m).cpu().numpy()
        print(f"member {m} ({c['Nx']}x{c['Ny']}x{c['Nz']}, k={c['k']}): "
              f"max |diff| {np.abs(got - want).max():.3e}", flush=True)
        np.testing.assert_array_equal(got, want, err_msg=str(m))


def test_members_on_every_sweep_tier(gpu):
    """The refinement sweep runs one launch per LDS tier of system sizes
    (profiles staged with two blocks per CU, gathered with two, staged with
    one): members on every tier, side by side, still equal their solo runs
    bit for bit, replayed."""
    from suzerain_amd._lib import load
    from suzerain_amd.channel import DEFAULT_SCENARIO
    from suzerain_amd.ensemble import Ensemble, default_dt
    lib = load()
    t0, t1 = lib.sz_internal_imexop_sweep_tier_n(0), lib.sz_internal_imexop_sweep_tier_n(1)
    sc = dict(DEFAULT_SCENARIO, Re=2000.0, Ma=0.8)
    members = [dict(Nx=4, Ny=t1 + 6, Nz=4, k=8, htdelta=3.0, scenario=sc),
               _members()[0],
               dict(Nx=8, Ny=t0 + 6, Nz=4, k=6, htdelta=3.0, scenario=sc),
               dict(Nx=4, Ny=t0, Nz=8, k=8, htdelta=3.0, scenario=sc)]
    print(f"case sweep tiers: bounds {t0}, {t1}; member n {[c['Ny'] for c in members]}",
          flush=True)
    assert t0 < t0 + 6 <= t1 < t1 + 6
    W = [_initial(c, gpu) for c in members]
    dts = [default_dt(c) for c in members]
    ens = Ensemble(members, dt=dts, solver=SPEC, graph=True)
    for m in range(len(members)):
        ens.set_state(m, W[m])
    steps = 2
    ens.step(steps)
    ens.check()
    for m, c in enumerate(members):
        want = _alone(c, W[m], dts[m], steps, gpu)
        got = ens.get_state(m).cpu().numpy()
        print(f"member {m} (n={c['Ny']}, k={c['k']}): max |diff| "
              f"{np.abs(got - want).max():.3e}", flush=True)
        np.testing.assert_array_equal(got, want, err_msg=str(m))


def test_autotune_leaves_states(gpu):
    from suzerain_amd.ensemble import Ensemble
    members = _members()
    W = [_initial(c, gpu) for c in members]
    dts = [2e-3, 1.5e-3, 2.5e-3, 1e-3]
    a = Ensemble(members, dt=dts, solver=SPEC, graph=True, max_streams=4)
    b = Ensemble(members, dt=dts, solver=SPEC, graph=True, max_streams=4)
    for m in range(len(members)):
        a.set_state(m, W[m])
        b.set_state(m, W[m])
    tune = a.autotune(steps=1)
    print(f"autotune: {tune}")
    assert len(tune["trial_ms"]) == 8 and tune["nstreams"] in (1, 2, 3, 4)
    assert tune["lu_bs"] in (4, 8)
    a.step(2)
    b.step(2)
    for m in range(len(members)):
        np.testing.assert_array_equal(a.get_state(m).cpu().numpy(),
                                      b.get_state(m).cpu().numpy())
