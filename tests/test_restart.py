"""Restart files on the host (no GPU): the library's HDF5 writer and reader
(csrc/sz_restart.cpp) against the reference's own restart file and against
the independent Python reader of tests/golden/hdf5_read.py.

* the library reads the reference's apps/perfect/manual.gold/
  channel.rhome_xyz/initial.h5 (copied to tests/golden/) and finds the state
  the Python reader transcribed, bit for bit, and the run's specification;
* files the library writes (float64, int32 and complex datasets of several
  ranks, names in any order) read back identically through both readers.
"""
import ctypes as C
import json
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import hdf5_read  # noqa: E402

GOLD_H5 = os.path.join(HERE, "golden", "channel_rhome_xyz_initial.h5")
KINDS = {0: np.float64, 1: np.int32, 2: np.complex128}


def _lib():
    from suzerain_amd._lib import load
    return load()


def lib_read(path, name):
    lib = _lib()
    f = lib.sz_internal_h5_read
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_int, C.c_void_p, C.c_longlong, C.c_char_p, C.c_int]
    kind, rank = C.c_int(), C.c_int()
    dims = (C.c_longlong * 8)()
    err = C.create_string_buffer(256)
    assert f(path.encode(), name.encode(), C.byref(kind), C.byref(rank), dims, 8,
             None, 0, err, 256) == 0, err.value
    shape = tuple(dims[i] for i in range(rank.value))
    out = np.zeros(shape, dtype=KINDS[kind.value])
    assert f(path.encode(), name.encode(), C.byref(kind), C.byref(rank), dims, 8,
             out.ctypes.data, out.nbytes, err, 256) == 0, err.value
    return out


def lib_write(path, sets):
    lib = _lib()
    f = lib.sz_internal_h5_write
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
    names = (C.c_char_p * len(sets))(*[n.encode() for n, _ in sets])
    kinds = (C.c_int * len(sets))(*[{np.float64: 0, np.int32: 1, np.complex128: 2}[a.dtype.type]
                                    for _, a in sets])
    ranks = (C.c_int * len(sets))(*[a.ndim for _, a in sets])
    dims = [d for _, a in sets for d in a.shape]
    cdims = (C.c_longlong * max(1, len(dims)))(*dims)
    arrays = [np.ascontiguousarray(a) for _, a in sets]
    data = (C.c_void_p * len(sets))(*[a.ctypes.data for a in arrays])
    err = C.create_string_buffer(256)
    assert f(path.encode(), len(sets), names, kinds, ranks, cdims, data, err, 256) == 0, err.value


def py_read(path, name):
    f = hdf5_read.File(path)
    arr, _ = f.read(f.walk()["/" + name])
    if arr.ndim and arr.shape[-1] == 2 and arr.dtype == np.float64 and name.startswith("rho"):
        arr = arr[..., 0] + 1j * arr[..., 1]
    return arr


def test_library_reads_reference_restart():
    with open(os.path.join(HERE, "golden", "channel_rhome_xyz_gold.json")) as fh:
        ini = json.load(fh)["initial"]
    for name in ("rho_E", "rho_u", "rho_v", "rho_w", "rho"):
        got = lib_read(GOLD_H5, name)
        s = ini["state"][name]
        want = (np.array(s["re"]) + 1j * np.array(s["im"])).reshape(s["shape"])
        assert got.shape == tuple(s["shape"]), (name, got.shape)
        np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(lib_read(GOLD_H5, "breakpoints_y"), ini["breakpoints_y"])
    assert lib_read(GOLD_H5, "Ny")[0] == 32
    assert lib_read(GOLD_H5, "t")[0] == 0.0


def test_restart_spec_of_reference_file():
    from suzerain_amd.channel import restart_spec
    sp = restart_spec(GOLD_H5)
    print(sp)
    assert (sp["Nx"], sp["Ny"], sp["Nz"], sp["k"]) == (1, 32, 1, 8)
    assert sp["htdelta"] == 2.0 and sp["Ly"] == 2.0
    assert sp["scenario"]["Re"] == 100.0 and sp["scenario"]["Ma"] == 1.5
    assert sp["bulk"] == {"rho_u": 1.0, "rho": 1.0}
    assert sp["lower"] == (1.0, 0.0, 0.0, 0.0)


@pytest.mark.parametrize("nsets", [1, 9, 40])
def test_written_files_read_back(tmp_path, nsets):
    rng = np.random.default_rng(nsets)
    sets = []
    for i in range(nsets):
        kind = i % 3
        shape = [(1,), (7,), (3, 4), (2, 3, 5)][i % 4]
        if kind == 0:
            a = rng.standard_normal(shape)
        elif kind == 1:
            a = rng.integers(-1000, 1000, shape).astype(np.int32)
        else:
            a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        sets.append((f"d{(7 * i) % nsets:03d}_{kind}", a))
    path = str(tmp_path / "x.h5")
    lib_write(path, sets)
    f = hdf5_read.File(path)
    assert sorted(f.walk()) == sorted("/" + n for n, _ in sets)
    for name, a in sets:
        np.testing.assert_array_equal(lib_read(path, name), a)
        arr, _ = f.read(f.walk()["/" + name])
        if a.dtype == np.complex128:
            arr = arr[..., 0] + 1j * arr[..., 1]
        np.testing.assert_array_equal(arr, a)
