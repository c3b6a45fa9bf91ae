"""Transcribe the reference's gold channel states into a JSON fixture.

Reads apps/perfect/manual.gold/channel.rhome_xyz/{initial,restart00000}.h5
(HDF5 data files the reference's regression check compares against,
apps/perfect/manual.gold/channel.rhome_xyz/check.sh) with the minimal reader
in hdf5_read.py and writes channel_rhome_xyz_gold.json: the grid and
scenario, the B-spline breakpoints, collocation points, integration
weights and Gram (Galerkin L2) band, the initial state and the state after advancing to t = 0.000919
with the rhome_xyz hybrid implicit operator, and the running means of the
implicit (bulk constraint) forcing saved with the restart; and copies
initial.h5 (a data file) to channel_rhome_xyz_initial.h5 for the restart
reader's tests.

Usage: python tests/golden/make_gold_state.py [/root/reference]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import hdf5_read  # noqa: E402

CASE = os.path.join("apps", "perfect", "manual.gold", "channel.rhome_xyz")
FIELDS = ("rho_E", "rho_u", "rho_v", "rho_w", "rho")
SCALARS = ("Nx", "Ny", "Nz", "k", "htdelta", "Lx", "Ly", "Lz", "DAFx", "DAFz",
           "Re", "Ma", "Pr", "gamma", "beta", "alpha", "bulk_rho", "bulk_rho_u",
           "bulk_rho_E", "evmagfactor", "lower_T", "lower_u", "lower_v",
           "lower_w", "upper_T", "upper_u", "upper_v", "upper_w", "t")
MEANS = ("bar_f", "bar_f_dot_u", "bar_qb", "bar_CrhoE", "bar_C2rhoE", "bar_Crhou",
         "bar_C2rhou", "bar_Crho", "bar_C2rho", "bar_Crhou_dot_u",
         "bar_C2rhou_dot_u")


def _num(v):
    v = float(v)
    return "nan" if v != v else v


def read_case(path):
    f = hdf5_read.File(path)
    paths = f.walk()

    def get(name):
        arr, _ = f.read(paths["/" + name])
        return arr

    out = {}
    for s in SCALARS:
        a = get(s).ravel()[0]
        out[s] = int(a) if s in ("Nx", "Ny", "Nz", "k") else _num(a)
    for s in ("breakpoints_y", "collocation_points_y", "integration_weights"):
        out[s] = [float(v) for v in get(s).ravel()]
    # the Galerkin L2 (Gram) operator, transposed band storage rows
    g = get("Gy0T")
    out["Gy0T"] = {"shape": list(g.shape), "values": [float(v) for v in g.ravel()]}
    # state fields: (Nz, Nx/2+1, Ny) complex stored as trailing (re, im)
    state = {}
    for name in FIELDS:
        a = get(name)
        state[name] = {"shape": list(a.shape[:-1]),
                       "re": [float(v) for v in a[..., 0].ravel()],
                       "im": [float(v) for v in a[..., 1].ravel()]}
    out["state"] = state
    means = {}
    for name in MEANS:
        a = get(name)
        means[name[4:]] = {"shape": list(a.shape),
                           "values": [_num(v) for v in a.ravel()]}
    out["implicit_means"] = means
    return out


def main(ref):
    base = os.path.join(ref, CASE)
    # the initial restart file itself, for the restart reader's tests
    import shutil
    shutil.copyfile(os.path.join(base, "initial.h5"),
                    os.path.join(HERE, "channel_rhome_xyz_initial.h5"))
    data = {
        "source": CASE + "/{initial,restart00000}.h5",
        "note": "one channel advanced from t = 0 to 0.000919 by perfect_advance "
                "--implicit=rhome_xyz (check.sh); states are B-spline "
                "coefficients (Nz, Nx/2+1, Ny)",
        "initial": read_case(os.path.join(base, "initial.h5")),
        "restart": read_case(os.path.join(base, "restart00000.h5")),
    }
    with open(os.path.join(HERE, "channel_rhome_xyz_gold.json"), "w") as fh:
        json.dump(data, fh, indent=0)
        fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
