"""Restart files with the GPU channel (the reference's ESIO/HDF5 layout).

* the reference's initial.h5 (apps/perfect/manual.gold/channel.rhome_xyz)
  loads into a channel built from its own specification (restart_spec) and
  one step reproduces the reference's advanced state;
* a 96^3 state saved and loaded into a fresh channel comes back bit for bit
  with its time, and the file's fields, read by the independent Python
  reader, are the retained modes of the state;
* an ensemble member checkpoints through the same call: its saved state,
  loaded into a channel of its own, steps to the member's next state per
  field <= 1e-12.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD_H5 = os.path.join(HERE, "golden", "channel_rhome_xyz_initial.h5")


def test_reference_restart_loads_and_steps(gpu):
    from suzerain_amd.channel import Channel, restart_spec
    from test_oracle_gold import BARS, field_errors, gold_state, load_gold
    print("case load reference initial.h5", flush=True)
    gold = load_gold()
    spec = restart_spec(GOLD_H5)
    ch = Channel(**spec)
    ch.load(GOLD_H5)
    got = ch.get_state(0).cpu().numpy()
    np.testing.assert_array_equal(got[:, 0, 0], gold_state(gold["initial"]))
    rest = got.copy()
    rest[:, 0, 0] = 0
    assert not rest.any()
    assert ch.time == 0.0
    dt = ch.step(max_dt=gold["restart"]["t"])
    errs = field_errors(ch.get_state(0).cpu().numpy()[:, 0, 0],
                        gold_state(gold["restart"]))
    print(f"after one step vs reference restart {errs}")
    for f in range(5):
        assert errs[f] <= BARS[f], errs
    assert ch.time == dt


def test_save_load_round_trip(gpu, tmp_path):
    import torch
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import hdf5_read
    from suzerain_amd.channel import Channel, laminar_state, restart_spec
    print("case 96^3 save/load", flush=True)
    cfg = dict(Nx=96, Ny=96, Nz=96, dNx=144, dNz=144, k=8, htdelta=3.0,
               bulk={"rho": 1.0, "rho_u": 1.0})
    a = Channel(**cfg)
    W = torch.from_numpy(laminar_state(a)).to(gpu)
    a.set_state(W)
    a.step(max_dt=0.01)
    a.step(max_dt=0.01)
    path = str(tmp_path / "restart.h5")
    a.save(path)
    spec = restart_spec(path)
    assert (spec["Nx"], spec["Ny"], spec["Nz"], spec["dNx"], spec["dNz"]) == (96, 96, 96, 144, 144)
    assert spec["bulk"] == {"rho": 1.0, "rho_u": 1.0}
    b = Channel(**spec)
    b.load(path)
    sa, sb = a.get_state(0).cpu().numpy(), b.get_state(0).cpu().numpy()
    np.testing.assert_array_equal(sa, sb)
    assert b.time == a.time == 0.02
    f = hdf5_read.File(path)
    arr, _ = f.read(f.walk()["/rho_u"])
    field = arr[..., 0] + 1j * arr[..., 1]                 # (Nz, Nx/2+1, Ny)
    assert field.shape == (96, 49, 96)
    for n in (0, 1, 47, 48, 49, 95):
        wz = n if n < 49 else n - 96
        cz = wz if wz >= 0 else 144 + wz
        want = sa[1, cz, :49] if abs(wz) <= 47 else np.zeros((49, 96))
        want = want.copy()
        want[48] = 0                                       # x Nyquist not kept
        np.testing.assert_array_equal(field[n], want)
    # and the two channels step on identically
    a.step(max_dt=0.01)
    b.step(max_dt=0.01)
    np.testing.assert_array_equal(a.get_state(0).cpu().numpy(),
                                  b.get_state(0).cpu().numpy())


def test_ensemble_member_checkpoint(gpu, tmp_path):
    import torch
    from suzerain_amd.channel import DEFAULT_SCENARIO, Channel, laminar_state
    from suzerain_amd.ensemble import Ensemble
    print("case ensemble member checkpoint", flush=True)
    spec = "zcgbsvx,reuse=true,siter=3,diter=5,tolsc=0.1"
    cfgs = [dict(Nx=8, Ny=18, Nz=12, k=8, htdelta=3.0,
                 scenario=dict(DEFAULT_SCENARIO, Re=1500.0, Ma=0.6)),
            dict(Nx=12, Ny=14, Nz=8, k=6, htdelta=3.0,
                 scenario=dict(DEFAULT_SCENARIO, Re=4000.0, Ma=1.2))]
    dts = [2e-3, 1.5e-3]
    ens = Ensemble(cfgs, dt=dts, solver=spec, graph=False)
    for m, c in enumerate(cfgs):
        W = laminar_state(Channel(**c, solver="zgbsv"), seed=11 + m)
        ens.set_state(m, torch.from_numpy(W).to(gpu))
    ens.step(2)
    path = str(tmp_path / "member0.h5")
    ens.member(0).save(path)
    ens.step(1)
    want = ens.get_state(0).cpu().numpy()
    solo = Channel(**cfgs[0], solver=spec)
    solo.load(path)
    assert solo.time == 2 * dts[0]
    solo.step(max_dt=dts[0])
    got = solo.get_state(0).cpu().numpy()
    errs = [float(np.abs(got[f] - want[f]).max() / np.abs(want[f]).max()) for f in range(5)]
    print(f"member 0 restarted alone vs in the ensemble: {errs}")
    assert max(errs) <= 1e-12, errs
