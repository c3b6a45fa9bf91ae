"""The reference's own regression states against the GPU channel.

apps/perfect/manual.gold/channel.rhome_xyz (tests/golden/
channel_rhome_xyz_gold.json): a Ny = 32, k = 8 channel with bulk rho and
rho u driven to 1, advanced by the reference's perfect_advance
--implicit=rhome_xyz to t = 0.000919.  One SMR91 step of the GPU channel
with the bulk constraint treatment, with the FP64 (zgbsv) and the mixed
precision (zcgbsvx) solvers, reproduces the reference's advanced state per
field <= 1e-12 of the field's magnitude (the wall-normal momentum, six
orders below the others, <= 1e-10 of its own); laminar_state without
perturbations is the reference's initial state.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PREC = "zcgbsvx,reuse=true,siter=3,diter=5,tolsc=0.1"


def _channel(gold, solver):
    from suzerain_amd.channel import Channel
    from test_oracle_gold import _f
    ini = gold["initial"]
    sc = {q: ini[q] for q in ("Re", "Pr", "Ma", "alpha", "beta", "gamma")}
    bulk = {"rho": _f(ini["bulk_rho"]), "rho_u": _f(ini["bulk_rho_u"])}
    if not math.isnan(_f(ini["bulk_rho_E"])):
        bulk["rho_E"] = _f(ini["bulk_rho_E"])
    lower = tuple(ini[f"lower_{q}"] for q in ("T", "u", "v", "w"))
    upper = tuple(ini[f"upper_{q}"] for q in ("T", "u", "v", "w"))
    # Nx = Nz = 1: the (0,0) mode alone; transforms of length 2 (any even
    # length gives the same mean-mode arithmetic)
    return Channel(Nx=ini["Nx"], Ny=ini["Ny"], Nz=ini["Nz"], dNx=2, dNz=2,
                   k=ini["k"], htdelta=ini["htdelta"], Lx=ini["Lx"],
                   Ly=ini["Ly"], Lz=ini["Lz"], scenario=sc, lower=lower,
                   upper=upper, solver=solver, bulk=bulk)


@pytest.mark.parametrize("solver", ["zgbsv", PREC])
def test_gpu_step_reproduces_reference(gpu, solver):
    import torch
    from test_oracle_gold import BARS, field_errors, gold_state, load_gold
    print(f"case gold channel.rhome_xyz solver={solver}", flush=True)
    gold = load_gold()
    ch = _channel(gold, solver)
    W0 = np.zeros(ch.shape, dtype=complex)
    W0[:, 0, 0] = gold_state(gold["initial"])
    ch.set_state(torch.from_numpy(W0).to(gpu))
    dt = ch.step(max_dt=gold["restart"]["t"])
    assert dt == gold["restart"]["t"]
    got = ch.get_state(0).cpu().numpy()
    errs = field_errors(got[:, 0, 0], gold_state(gold["restart"]))
    print(f"GPU step vs reference per field {errs}", flush=True)
    for f in range(5):
        assert errs[f] <= BARS[f], (f, errs)
    # nothing outside the mean mode
    rest = got.copy()
    rest[:, 0, 0] = 0
    assert not rest.any()


def test_laminar_state_is_reference_initial(gpu):
    from suzerain_amd.channel import laminar_state
    from test_oracle_gold import gold_state, load_gold
    gold = load_gold()
    ch = _channel(gold, "zgbsv")
    W = laminar_state(ch, amp=0.0)
    want = gold_state(gold["initial"])
    err = np.abs(W[:, 0, 0] - want).max() / np.abs(want).max()
    print(f"laminar_state vs reference initial: {err:.2e}")
    assert err <= 1e-14
