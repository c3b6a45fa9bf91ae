"""Ensembles of channels on one GPU: the Python view of sz_ensemble.

    members = uncertainty_sweep(48, seed=...)       # config dicts
    ens = Ensemble(members, graph=True)
    for m, c in enumerate(members): ens.set_state(m, laminar_state(...))
    ens.autotune(steps=1)
    ens.step(10)

Each member is a channel (suzerain_amd.channel.Channel arguments) with its
own fixed time step; the ensemble advances every member by SMR91 steps,
solving all members' implicit systems together in the grouped mixed
precision solver.  The reference advances one case per process
(apps/perfect/main_advance.cpp); a sweep over Re, Ma and resolution is many
such runs.
"""
import ctypes as C
import math

import numpy as np

from ._lib import load, check, sz_isothermal_bc
from .channel import (Channel, DEFAULT_SCENARIO, sz_channel_spec,
                      sz_ns_scenario, laminar_state)

__all__ = ["Ensemble", "uncertainty_sweep", "member_spec", "member_bytes"]


class sz_ensemble_opts(C.Structure):
    _fields_ = [("solver", C.c_char_p), ("nstreams", C.c_int),
                ("max_streams", C.c_int), ("graph", C.c_int),
                ("lu_bs", C.c_int)]


_SIGS = {
    "sz_ensemble_create": (C.c_void_p, [C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_char_p,
                                        C.c_int]),
    "sz_ensemble_free": (None, [C.c_void_p]),
    "sz_ensemble_add": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double,
                                  C.c_char_p, C.c_int]),
    "sz_ensemble_remove": (C.c_int, [C.c_void_p, C.c_int]),
    "sz_ensemble_find": (C.c_int, [C.c_void_p, C.c_int]),
    "sz_ensemble_size": (C.c_int, [C.c_void_p]),
    "sz_ensemble_member": (C.c_void_p, [C.c_void_p, C.c_int]),
    "sz_ensemble_bytes": (C.c_longlong, [C.c_void_p]),
    "sz_ensemble_step": (C.c_int, [C.c_void_p, C.c_int]),
    "sz_ensemble_check": (C.c_int, [C.c_void_p]),
    "sz_ensemble_solver_stats": (C.c_void_p, [C.c_void_p, C.c_int]),
    "sz_ensemble_autotune": (C.c_int, [C.c_void_p, C.c_int]),
    "sz_ensemble_tuning": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int, C.c_void_p]),
    "sz_channel_state": (C.c_void_p, [C.c_void_p, C.c_int]),
    "sz_channel_geometry": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
}


def _lib():
    lib = load()
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


GRID_CHOICES = (32, 48, 64, 96, 128)


def uncertainty_sweep(count=48, seed=20261203, n_range=(48, 256), ks=(6, 8),
                      grids=GRID_CHOICES, Re=(1000.0, 6000.0), Ma=(0.3, 1.5)):
    """Seeded members of an uncertainty sweep: B-spline coefficients
    n ~ U{48..256}, order k in {6, 8}, Nx, Nz each from {32, 48, 64, 96, 128}
    (3/2 dealiasing), Re ~ U[1000, 6000], Ma ~ U[0.3, 1.5]; the remaining
    parameters are the channel defaults (htdelta 3, Lx 4 pi, Ly 2, Lz 4 pi/3,
    Pr 0.7, gamma 1.4)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        k = int(rng.choice(ks))
        Nx = int(rng.choice(grids))
        Nz = int(rng.choice(grids))
        re = float(rng.uniform(*Re))
        ma = float(rng.uniform(*Ma))
        out.append(dict(Nx=Nx, Ny=n, Nz=Nz, k=k, htdelta=3.0,
                        scenario=dict(DEFAULT_SCENARIO, Re=re, Ma=ma)))
    return out


def member_spec(cfg):
    """sz_channel_spec of a member config dict (Channel's arguments)."""
    Nx, Ny, Nz, k = cfg["Nx"], cfg["Ny"], cfg["Nz"], cfg.get("k", 8)
    dNx = cfg.get("dNx") or 3 * Nx // 2
    dNz = cfg.get("dNz") or 3 * Nz // 2
    sc = dict(DEFAULT_SCENARIO, **cfg.get("scenario", {}))
    lower = cfg.get("lower", (1.0, 0.0, 0.0, 0.0))
    upper = cfg.get("upper", (1.0, 0.0, 0.0, 0.0))
    return sz_channel_spec(
        Nx, Ny, Nz, dNx, dNz, k, cfg.get("htdelta", 3.0),
        cfg.get("Lx", 4 * math.pi), cfg.get("Ly", 2.0),
        cfg.get("Lz", 4 * math.pi / 3),
        sz_ns_scenario(*[sc[q] for q in ("alpha", "beta", "gamma", "Ma", "Pr",
                                         "Re")]),
        sz_isothermal_bc(1, 1, *lower, *upper), b"none", 1)


def member_bytes(cfg, slots=3):
    """Device bytes a member keeps resident in an ensemble: two states,
    references, and its share of the grouped solver (float factors per slot,
    pivots, double vectors b, x, A x and the float right hand side)."""
    Nx, Ny, Nz, k = cfg["Nx"], cfg["Ny"], cfg["Nz"], cfg.get("k", 8)
    dNx, dNz = 3 * Nx // 2, 3 * Nz // 2
    Xw = dNx // 2 + 1
    nsys = ((Nx - 1) // 2 + 1) * (2 * ((Nz - 1) // 2) + 1)
    kl = k - 2
    KL = 5 * (kl + 1) - 1
    N = 5 * Ny
    F = dNz * Xw * Ny
    state = 2 * 5 * F * 16
    lu = slots * nsys * N * (3 * KL + 1) * 8 + slots * nsys * N * 4
    vec = nsys * N * (2 * 16 + 8)
    return state + lu + vec + nsys * 400


class Ensemble:
    def __init__(self, members, dt=None, solver="zcgbsvx,reuse=true,siter=3,"
                 "diter=5,tolsc=0.1", nstreams=1, max_streams=4, graph=True,
                 lu_bs=0, stream=None):
        import torch
        self.members = list(members)
        nm = len(self.members)
        specs = (sz_channel_spec * nm)(*[member_spec(c) for c in self.members])
        if dt is None:
            dt = [default_dt(c) for c in self.members]
        self.dt = list(dt)
        dts = (C.c_double * nm)(*self.dt)
        opts = sz_ensemble_opts(solver.encode(), nstreams, max_streams,
                                int(bool(graph)), lu_bs)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._stream = stream
        err = C.create_string_buffer(512)
        self._h = _lib().sz_ensemble_create(nm, specs, dts, C.byref(opts),
                                            C.c_void_p(stream), err, 512)
        if not self._h:
            raise MemoryError(err.value.decode())
        self.ids = list(range(nm))
        self.shapes = [self._shape(c) for c in self.members]

    @staticmethod
    def _shape(c):
        dNz = c.get("dNz") or 3 * c["Nz"] // 2
        dNx = c.get("dNx") or 3 * c["Nx"] // 2
        return (5, dNz, dNx // 2 + 1, c["Ny"])

    def add(self, cfg, dt=None):
        """Admit a member; returns its id (MemoryError when it does not fit)."""
        dt = default_dt(cfg) if dt is None else dt
        err = C.create_string_buffer(512)
        spec = member_spec(cfg)
        mid = _lib().sz_ensemble_add(self._h, C.byref(spec), dt, err, 512)
        if mid < 0:
            raise MemoryError(err.value.decode())
        self.members.append(cfg)
        self.dt.append(dt)
        self.ids.append(mid)
        self.shapes.append(self._shape(cfg))
        return mid

    def remove(self, mid):
        """Retire member id (frees its device storage)."""
        k = self.ids.index(mid)
        check(_lib().sz_ensemble_remove(self._h, mid), "sz_ensemble_remove")
        for lst in (self.members, self.dt, self.ids, self.shapes):
            del lst[k]

    def position(self, mid):
        return self.ids.index(mid)

    def __len__(self):
        return len(self.members)

    def bytes(self):
        return _lib().sz_ensemble_bytes(self._h)

    def _state_ptr(self, m, which=0):
        return _lib().sz_channel_state(_lib().sz_ensemble_member(self._h, m),
                                       which)

    def set_state(self, m, t):
        """Copy a device tensor shaped like member m's state into it."""
        import torch
        t = t.to(dtype=torch.complex128).contiguous()
        assert tuple(t.shape) == self.shapes[m], (t.shape, self.shapes[m])
        h = C.CDLL("libamdhip64.so")
        rc = h.hipMemcpyAsync(C.c_void_p(self._state_ptr(m)),
                              C.c_void_p(t.data_ptr()), C.c_size_t(t.numel() * 16),
                              3, C.c_void_p(self._stream))
        h.hipStreamSynchronize(C.c_void_p(self._stream))
        if rc:
            raise RuntimeError(f"hipMemcpyAsync failed: {rc}")

    def get_state(self, m):
        import torch
        out = torch.empty(self.shapes[m], dtype=torch.complex128, device="cuda:0")
        h = C.CDLL("libamdhip64.so")
        h.hipStreamSynchronize(C.c_void_p(self._stream))
        rc = h.hipMemcpyAsync(C.c_void_p(out.data_ptr()),
                              C.c_void_p(self._state_ptr(m)),
                              C.c_size_t(out.numel() * 16), 3,
                              C.c_void_p(self._stream))
        h.hipStreamSynchronize(C.c_void_p(self._stream))
        if rc:
            raise RuntimeError(f"hipMemcpyAsync failed: {rc}")
        return out

    def member(self, m):
        """Non-owning Channel view of member m (states, candidates)."""
        c = self.members[m]
        h = _lib().sz_ensemble_member(self._h, m)
        return Channel.view(h, c["Nx"], c["Ny"], c["Nz"], c.get("dNx"),
                            c.get("dNz"), c.get("k", 8), c.get("htdelta", 3.0),
                            c.get("Lx", 4 * math.pi), c.get("Ly", 2.0),
                            c.get("Lz", 4 * math.pi / 3), c.get("scenario"),
                            stream=self._stream)

    def step(self, nsteps=1):
        check(_lib().sz_ensemble_step(self._h, nsteps), "sz_ensemble_step")

    def check(self):
        rc = _lib().sz_ensemble_check(self._h)
        if rc:
            raise ArithmeticError(f"ensemble implicit solve: info {rc}")

    def autotune(self, steps=1):
        check(_lib().sz_ensemble_autotune(self._h, steps), "sz_ensemble_autotune")
        return self.tuning()

    def tuning(self):
        ns, bs, cnt = C.c_int(), C.c_int(), C.c_int()
        ms = np.zeros(16)
        _lib().sz_ensemble_tuning(self._h, C.byref(ns), C.byref(bs),
                                  ms.ctypes.data, 16, C.byref(cnt))
        return {"nstreams": ns.value, "lu_bs": bs.value,
                "trial_ms": [round(float(x), 3) for x in ms[:cnt.value]]}

    def solver_stats(self, which):
        from .solver import device_to_host
        k = {"siter": 0, "diter": 1, "tolsc": 2, "res": 3, "info": 6}[which]
        p = _lib().sz_ensemble_solver_stats(self._h, k)
        dtype = np.int32 if which in ("siter", "diter", "info") else np.float64
        n = sum(_nsys(c) for c in self.members)
        return device_to_host(p, n * np.dtype(dtype).itemsize, dtype)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib().sz_ensemble_free(self._h)
                self._h = None
        except Exception:
            pass


def _nsys(cfg):
    return ((cfg["Nx"] - 1) // 2 + 1) * (2 * ((cfg["Nz"] - 1) // 2) + 1)


def default_dt(cfg, cfl=0.5):
    """A fixed time step for a member: cfl times the SMR91 stability limit of
    the fastest acoustic wave on the finest wall-normal spacing, estimated
    from the grid alone (the member's own step candidates refine it)."""
    Ny, k = cfg["Ny"], cfg.get("k", 8)
    Ma = cfg.get("scenario", {}).get("Ma", DEFAULT_SCENARIO["Ma"])
    htdelta, Ly = cfg.get("htdelta", 3.0), cfg.get("Ly", 2.0)
    # finest breakpoint spacing of the two-sided tanh stretching
    x = np.linspace(0.0, 1.0, Ny - k + 2)
    y = 0.5 * (1 + np.tanh(htdelta * (x - 0.5)) / np.tanh(htdelta / 2.0)) * Ly
    h = float(np.min(np.diff(y)))
    Nx, Nz = cfg["Nx"], cfg["Nz"]
    Lx, Lz = cfg.get("Lx", 4 * math.pi), cfg.get("Lz", 4 * math.pi / 3)
    c = 1.0 + 1.0 / Ma                    # |u| + sound speed, nondimensional
    lam = c * (math.pi / h / 3.0 + math.pi * Nx / Lx + math.pi * Nz / Lz)
    return cfl * 2.51 / lam
