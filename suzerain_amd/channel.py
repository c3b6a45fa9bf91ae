"""Channel solver on one GPU: Python view of the sz_channel C handle.

    ch = Channel(Nx=96, Ny=96, Nz=96, dNx=144, dNz=144, k=8, htdelta=3.0)
    ch.set_state(laminar_state(ch, seed=...))
    dt = ch.step()

The step is suzerain's SMR91 low-storage hybrid implicit/explicit step
(suzerain/lowstorage.hpp:1461-1520) with the rhome_xyz linearization of
apps/perfect (navier_stokes.hpp, operator_hybrid_isothermal.cpp).  State
arrays are complex128 (5, dNz, dNx/2+1, Ny): (rho E, rho u, rho v, rho w,
rho), y holding B-spline coefficients.
"""
import ctypes as C
import math

import numpy as np

from ._lib import load, check, sz_isothermal_bc

__all__ = ["Channel", "laminar_state", "restart_spec", "DEFAULT_SCENARIO", "SMR91"]

# SMR91 alpha, beta, gamma, zeta (suzerain/lowstorage.hpp:1241-1305)
SMR91 = ((29 / 96, -3 / 40, 1 / 6), (37 / 160, 5 / 24, 1 / 6),
         (8 / 15, 5 / 12, 3 / 4), (0.0, -17 / 60, -5 / 12))

# apps/perfect/main_initial.cpp:124-134 defaults with the channel case values
DEFAULT_SCENARIO = {"alpha": 0.0, "beta": 2.0 / 3.0, "gamma": 1.4, "Ma": 1.5,
                    "Pr": 0.7, "Re": 3000.0}


class sz_ns_scenario(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("alpha", "beta", "gamma", "Ma",
                                          "Pr", "Re")]


class sz_channel_spec(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
                ("dNx", C.c_int), ("dNz", C.c_int), ("k", C.c_int),
                ("htdelta", C.c_double), ("Lx", C.c_double),
                ("Ly", C.c_double), ("Lz", C.c_double),
                ("scenario", sz_ns_scenario), ("bc", sz_isothermal_bc),
                ("solver", C.c_char_p), ("linearize", C.c_int),
                ("constrained", C.c_int), ("bulk", C.c_double * 5)]


# state fields a bulk target may name (the reference's bulk_rho_E,
# bulk_rho_u, bulk_rho; rho_v and rho_w complete the treatment's set)
BULK_FIELDS = ("rho_E", "rho_u", "rho_v", "rho_w", "rho")
IMPLICIT_MEANS = ("fx fy fz f_dot_u qb CrhoE C2rhoE Crhou C2rhou Crhov C2rhov "
                  "Crhow C2rhow Crho C2rho Crhou_dot_u C2rhou_dot_u").split()


def bulk_spec(bulk):
    """(constrained, 5 targets) from a dict naming BULK_FIELDS (others
    NaN: disabled), or (0, zeros) for None."""
    if not bulk:
        return 0, (C.c_double * 5)()
    unknown = set(bulk) - set(BULK_FIELDS)
    if unknown:
        raise ValueError(f"unknown bulk fields {sorted(unknown)}")
    return 1, (C.c_double * 5)(*[float(bulk.get(f, math.nan))
                                 for f in BULK_FIELDS])


_SIGS = {
    "sz_channel_create": (C.c_void_p, [C.c_void_p, C.c_void_p, C.c_char_p,
                                       C.c_int]),
    "sz_channel_free": (None, [C.c_void_p]),
    "sz_channel_geometry": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    "sz_channel_state": (C.c_void_p, [C.c_void_p, C.c_int]),
    "sz_channel_buffer": (C.c_void_p, [C.c_void_p, C.c_int]),
    "sz_channel_nonlinear": (C.c_int, [C.c_void_p, C.c_int, C.c_int,
                                       C.c_void_p]),
    "sz_channel_accumulate_operator": (C.c_int, [C.c_void_p, C.c_double,
                                                 C.c_int, C.c_double,
                                                 C.c_int]),
    "sz_channel_apply_operator": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "sz_channel_invert_operator": (C.c_int, [C.c_void_p, C.c_double,
                                             C.c_int]),
    "sz_channel_invert_operator_slot": (C.c_int, [C.c_void_p, C.c_double,
                                                  C.c_int, C.c_int]),
    "sz_channel_invert_operator_substep": (C.c_int, [C.c_void_p, C.c_double,
                                                     C.c_int, C.c_int,
                                                     C.c_double, C.c_int]),
    "sz_channel_solver_slots": (C.c_int, [C.c_void_p]),
    "sz_channel_solver": (C.c_void_p, [C.c_void_p, C.c_int]),
    "sz_channel_solver_stats": (C.c_void_p, [C.c_void_p, C.c_int]),
    "sz_channel_solver_counts": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "sz_channel_check": (C.c_int, [C.c_void_p]),
    "sz_channel_add_scaled": (C.c_int, [C.c_void_p, C.c_double, C.c_int,
                                        C.c_int]),
    "sz_channel_copy": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "sz_channel_exchange": (None, [C.c_void_p]),
    "sz_channel_step": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p]),
    "sz_channel_set_manufactured": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sz_channel_set_time": (C.c_int, [C.c_void_p, C.c_double]),
    "sz_channel_time": (C.c_double, [C.c_void_p]),
    "sz_channel_nonlinear_time": (C.c_int, [C.c_void_p, C.c_int, C.c_int,
                                            C.c_double, C.c_void_p]),
    "sz_channel_manufactured_state": (C.c_int, [C.c_void_p, C.c_int,
                                                C.c_double]),
    "sz_channel_field_L2": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p,
                                      C.c_void_p]),
    "sz_channel_save": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p,
                                  C.c_int]),
    "sz_channel_load": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p,
                                  C.c_int]),
    "sz_restart_read_spec": (C.c_int, [C.c_char_p, C.c_void_p, C.c_char_p,
                                       C.c_int]),
    "sz_channel_factor_timing": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_void_p]),
    "sz_channel_substep_timing": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p,
                                            C.c_int, C.c_void_p]),
}


def _lib():
    lib = load()
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


def _hip():
    return C.CDLL("libamdhip64.so")


class Channel:
    def __init__(self, Nx=96, Ny=96, Nz=96, dNx=None, dNz=None, k=8,
                 htdelta=3.0, Lx=4 * math.pi, Ly=2.0, Lz=4 * math.pi / 3,
                 scenario=None, lower=(1.0, 0.0, 0.0, 0.0),
                 upper=(1.0, 0.0, 0.0, 0.0), solver="zgbsv", stream=None,
                 bulk=None):
        """bulk: bulk (y-averaged) targets of the (0,0) mode by field name
        (BULK_FIELDS), e.g. {"rho": 1, "rho_u": 1, "rho_E": e_b}; the
        implicit solve then applies the integral constraint treatment."""
        import torch
        sc = dict(DEFAULT_SCENARIO, **(scenario or {}))
        self.scenario = sc
        dNx = dNx or 3 * Nx // 2
        dNz = dNz or 3 * Nz // 2
        self.Nx, self.Ny, self.Nz, self.dNx, self.dNz, self.k = (
            Nx, Ny, Nz, dNx, dNz, k)
        self.htdelta, self.Lx, self.Ly, self.Lz = htdelta, Lx, Ly, Lz
        self.lower, self.upper = lower, upper
        spec = sz_channel_spec(
            Nx, Ny, Nz, dNx, dNz, k, htdelta, Lx, Ly, Lz,
            sz_ns_scenario(*[sc[q] for q in ("alpha", "beta", "gamma", "Ma",
                                             "Pr", "Re")]),
            sz_isothermal_bc(1, 1, *lower, *upper), solver.encode(), 1,
            *bulk_spec(bulk))
        self.bulk = dict(bulk or {})
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._stream = stream
        err = C.create_string_buffer(512)
        self._h = _lib().sz_channel_create(C.byref(spec), C.c_void_p(stream),
                                           err, 512)
        if not self._h:
            raise ValueError(err.value.decode())
        Xw, nsys = C.c_int(), C.c_int()
        F, Fp = C.c_longlong(), C.c_longlong()
        _lib().sz_channel_geometry(self._h, C.byref(Xw), C.byref(nsys),
                                   C.byref(F), C.byref(Fp))
        self.Xw, self.nsys, self.F, self.Fp = Xw.value, nsys.value, F.value, Fp.value
        self.shape = (5, dNz, self.Xw, Ny)

    @classmethod
    def view(cls, handle, Nx, Ny, Nz, dNx=None, dNz=None, k=8, htdelta=3.0,
             Lx=4 * math.pi, Ly=2.0, Lz=4 * math.pi / 3, scenario=None,
             lower=(1.0, 0.0, 0.0, 0.0), upper=(1.0, 0.0, 0.0, 0.0),
             stream=None):
        """A non-owning Channel over an existing sz_channel handle (an
        ensemble member) made from the same arguments."""
        import torch
        self = cls.__new__(cls)
        self.scenario = dict(DEFAULT_SCENARIO, **(scenario or {}))
        dNx = dNx or 3 * Nx // 2
        dNz = dNz or 3 * Nz // 2
        self.Nx, self.Ny, self.Nz, self.dNx, self.dNz, self.k = (
            Nx, Ny, Nz, dNx, dNz, k)
        self.htdelta, self.Lx, self.Ly, self.Lz = htdelta, Lx, Ly, Lz
        self.lower, self.upper = lower, upper
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._stream = stream
        self._h = handle
        self._owned = False
        Xw, nsys = C.c_int(), C.c_int()
        F, Fp = C.c_longlong(), C.c_longlong()
        _lib().sz_channel_geometry(self._h, C.byref(Xw), C.byref(nsys),
                                   C.byref(F), C.byref(Fp))
        self.Xw, self.nsys, self.F, self.Fp = Xw.value, nsys.value, F.value, Fp.value
        self.shape = (5, dNz, self.Xw, Ny)
        return self

    def candidates(self):
        """The 12 stable step candidates of the last zeroth substep."""
        out = np.empty(12)
        p = _lib().sz_channel_buffer(self._h, 10)
        self._copy(out.ctypes.data, p, out.nbytes, 2)
        return out

    # -------------------------------------------------------------- access
    def _copy(self, dst, src, nbytes, kind):
        rc = _hip().hipMemcpyAsync(C.c_void_p(dst), C.c_void_p(src),
                                   C.c_size_t(nbytes), kind,
                                   C.c_void_p(self._stream))
        if rc:
            raise RuntimeError(f"hipMemcpyAsync failed: {rc}")
        _hip().hipStreamSynchronize(C.c_void_p(self._stream))

    def state_ptr(self, which=0):
        return _lib().sz_channel_state(self._h, which)

    def set_state(self, t, which=0):
        """Copy a device tensor (5, dNz, Xw, Ny) complex128 into a state."""
        import torch
        t = t.to(dtype=torch.complex128).contiguous()
        assert tuple(t.shape) == self.shape
        self._copy(self.state_ptr(which), t.data_ptr(), t.numel() * 16, 3)

    def get_state(self, which=0):
        import torch
        out = torch.empty(self.shape, dtype=torch.complex128, device="cuda:0")
        self._copy(out.data_ptr(), self.state_ptr(which), out.numel() * 16, 3)
        return out

    def refs(self):
        """Plane means (42, Ny) of the last zeroth-substep collection."""
        out = np.empty((42, self.Ny))
        p = _lib().sz_channel_buffer(self._h, 0)
        self._copy(out.ctypes.data, p, out.nbytes, 2)
        return out

    def _host(self, which, shape):
        p = _lib().sz_channel_buffer(self._h, which)
        if not p:
            return None
        out = np.empty(shape)
        self._copy(out.ctypes.data, p, out.nbytes, 2)
        return out

    def implicit_means(self):
        """Running means of the implicit forcing (17, Ny) by
        IMPLICIT_MEANS name order; None without bulk constraints."""
        return self._host(11, (17, self.Ny))

    def constraint_multipliers(self):
        """The 5 multipliers (per field) of the last constraint solve."""
        return self._host(12, (5,))

    def bulk_weights(self):
        """w with w . c = (1/Ly) integral over y of the B-spline
        expansion c."""
        return self._host(13, (self.Ny,))

    def instantaneous(self):
        """Plane means (42, Ny) of the last non-zeroth nonlinear
        evaluation (kept with bulk constraints)."""
        return self._host(14, (42, self.Ny))

    def collocation_points(self):
        w = _lib().sz_channel_buffer(self._h, 8)
        f = load().sz_bsplineop_collocation_point
        return np.array([f(w, i) for i in range(self.Ny)])

    def operators(self):
        """Host copies (3, Ny, ld) of the transposed bands of M, D1, D2 and
        (kl, ku, max_kl, max_ku)."""
        lib = load()
        w = _lib().sz_channel_buffer(self._h, 8)
        ld, n = lib.sz_bsplineop_ld(w), self.Ny
        dt = np.stack([np.ctypeslib.as_array(lib.sz_bsplineop_D_T(w, d),
                                             shape=(n * ld,)).copy().reshape(n, ld)
                       for d in range(3)])
        kl = [lib.sz_bsplineop_kl(w, d) for d in range(3)]
        ku = [lib.sz_bsplineop_ku(w, d) for d in range(3)]
        return dt, kl, ku, lib.sz_bsplineop_max_kl(w), lib.sz_bsplineop_max_ku(w)

    def lambdas(self):
        out = np.empty(2 * self.Ny)
        p = _lib().sz_channel_buffer(self._h, 7)
        self._copy(out.ctypes.data, p, out.nbytes, 2)
        return out[:self.Ny], out[self.Ny:]

    # ----------------------------------------------------------- operators
    def nonlinear(self, which=1, zeroth=True):
        cand = np.zeros(12)
        rc = _lib().sz_channel_nonlinear(self._h, which, int(zeroth),
                                         cand.ctypes.data if zeroth else None)
        check(rc, "sz_channel_nonlinear")
        return cand if zeroth else None

    def accumulate_operator(self, phi, src, beta, dst):
        check(_lib().sz_channel_accumulate_operator(self._h, phi, src, beta, dst),
              "sz_channel_accumulate_operator")

    def apply_operator(self, phi, which=0):
        check(_lib().sz_channel_apply_operator(self._h, phi, which),
              "sz_channel_apply_operator")

    def invert_operator(self, phi, which=0, slot=0, delta_t=0.0, substep=0):
        """delta_t > 0: a substep of a step of that size (the constraint
        treatment's running means)."""
        check(_lib().sz_channel_invert_operator_substep(self._h, phi, which, slot,
                                                        delta_t, substep),
              "sz_channel_invert_operator_substep")
        rc = _lib().sz_channel_check(self._h)
        if rc:
            raise ArithmeticError(f"implicit solve: zero pivot, info {rc}")

    def solver_slots(self):
        """Factorization slots: 3 for zcgbsvx with reuse (one per substep)."""
        return _lib().sz_channel_solver_slots(self._h)

    def solver_stats(self, which):
        """Per-system outcome of the last mixed precision (zcgbsvx) solve:
        "siter", "diter", "tolsc", "res", "fact", "apprx" or "info" (None
        for zgbsv)."""
        from .solver import device_to_host
        k = {"siter": 0, "diter": 1, "tolsc": 2, "res": 3, "fact": 4,
             "apprx": 5, "info": 6}[which]
        p = _lib().sz_channel_solver_stats(self._h, k)
        if not p:
            return None
        dtype = {"siter": np.int32, "diter": np.int32, "tolsc": np.float64,
                 "res": np.float64, "fact": np.int8, "apprx": np.int32,
                 "info": np.int32}[which]
        a = device_to_host(p, self.nsys * np.dtype(dtype).itemsize, dtype)
        return a.view("S1") if which == "fact" else a

    def solver_counts(self):
        """(systems factored, restarted, sent to double precision) by the
        last mixed precision solve."""
        f, r, d = C.c_int(0), C.c_int(0), C.c_int(0)
        rc = _lib().sz_channel_solver_counts(self._h, C.byref(f), C.byref(r),
                                             C.byref(d))
        return None if rc else (f.value, r.value, d.value)

    def add_scaled(self, alpha, src, dst):
        check(_lib().sz_channel_add_scaled(self._h, alpha, src, dst),
              "sz_channel_add_scaled")

    def exchange(self):
        _lib().sz_channel_exchange(self._h)

    def copy(self, src, dst):
        check(_lib().sz_channel_copy(self._h, src, dst), "sz_channel_copy")

    def step(self, max_dt=0.0):
        dt = C.c_double(0)
        rc = _lib().sz_channel_step(self._h, max_dt, C.byref(dt))
        if rc > 0:
            raise ArithmeticError(f"implicit solve: zero pivot, info {rc}")
        check(rc, "sz_channel_step")
        return dt.value

    # ------------------------------------------------ manufactured solutions
    def set_manufactured(self, m):
        """m: a suzerain_amd.mms.Manufactured (its forcing joins N), or None."""
        self._mms = m
        check(_lib().sz_channel_set_manufactured(
            self._h, C.byref(m.s) if m is not None else None),
            "sz_channel_set_manufactured")

    @property
    def time(self):
        return _lib().sz_channel_time(self._h)

    def set_time(self, t):
        check(_lib().sz_channel_set_time(self._h, t), "sz_channel_set_time")

    def manufactured_state(self, t, which=0):
        """State `which` <- the manufactured solution at time t."""
        check(_lib().sz_channel_manufactured_state(self._h, which, t),
              "sz_channel_manufactured_state")

    def field_L2(self, which=0):
        """(mean, fluctuating) L2 norms over the domain per field."""
        mean, fluct = np.zeros(5), np.zeros(5)
        check(_lib().sz_channel_field_L2(self._h, which, mean.ctypes.data,
                                         fluct.ctypes.data), "sz_channel_field_L2")
        return mean, fluct

    # ------------------------------------------------------- restart files
    def save(self, path, which=0):
        """Write state `which` and the run's metadata in the reference's
        restart layout (HDF5)."""
        err = C.create_string_buffer(512)
        if _lib().sz_channel_save(self._h, which, str(path).encode(), err, 512):
            raise IOError(err.value.decode())

    def load(self, path, which=0):
        """Read a restart file's state into `which` (and the time)."""
        err = C.create_string_buffer(512)
        if _lib().sz_channel_load(self._h, which, str(path).encode(), err, 512):
            raise IOError(err.value.decode())

    def factor_timing(self, enable=-1):
        """(summed ms, count) of the factorization launches bracketed by HIP
        events since the last call; enable 1/0 switches bracketing on/off."""
        ms, n = C.c_double(0), C.c_int(0)
        _lib().sz_channel_factor_timing(self._h, enable, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def substep_timing(self, enable=-1, max_count=100000):
        """Device milliseconds of each substep recorded since the last call
        (numpy array); enable 1/0 switches recording on/off."""
        buf = np.zeros(max_count)
        n = C.c_int(0)
        check(_lib().sz_channel_substep_timing(self._h, enable,
                                               buf.ctypes.data, max_count,
                                               C.byref(n)),
              "sz_channel_substep_timing")
        return buf[:n.value].copy()

    def __del__(self):
        try:
            if getattr(self, "_h", None) and getattr(self, "_owned", True):
                _lib().sz_channel_free(self._h)
                self._h = None
        except Exception:
            pass


def restart_spec(path):
    """Channel(...) keyword arguments describing a restart file's run."""
    sp = sz_channel_spec()
    err = C.create_string_buffer(512)
    if _lib().sz_restart_read_spec(str(path).encode(), C.byref(sp), err, 512):
        raise IOError(err.value.decode())
    sc = {k: getattr(sp.scenario, k) for k in ("alpha", "beta", "gamma", "Ma",
                                               "Pr", "Re")}
    bc = sp.bc
    out = dict(Nx=sp.Nx, Ny=sp.Ny, Nz=sp.Nz, dNx=sp.dNx, dNz=sp.dNz, k=sp.k,
               htdelta=sp.htdelta, Lx=sp.Lx, Ly=sp.Ly, Lz=sp.Lz, scenario=sc,
               lower=(bc.lower_T, bc.lower_u, bc.lower_v, bc.lower_w),
               upper=(bc.upper_T, bc.upper_u, bc.upper_v, bc.upper_w))
    if sp.constrained:
        out["bulk"] = {f: sp.bulk[i] for i, f in enumerate(BULK_FIELDS)
                       if sp.bulk[i] == sp.bulk[i]}
    return out


def laminar_state(ch, seed=20261202, amp=1e-2, kmax=8):
    """Isothermal laminar channel (apps/perfect/main_initial.cpp:404-452:
    T = 1 walls, parabolic streamwise velocity with unit bulk velocity,
    constant density 1, p = rho T / gamma) plus seeded Fourier perturbations
    of the momentum of amplitude amp |k|^-2 for 0 < |k| <= kmax with a
    y(Ly - y) wall-normal shape, as B-spline coefficients on the host;
    returns a complex128 numpy array shaped ch.shape."""
    rng = np.random.default_rng(seed)
    sc = ch.scenario
    gamma, Ma = sc["gamma"], sc["Ma"]
    y = ch.collocation_points()
    Ly = ch.Ly
    dt, kl, ku, max_kl, max_ku = ch.operators()
    n = ch.Ny
    Mdense = np.zeros((n, n))
    for r in range(n):
        for c in range(max(0, r - ku[0]), min(n, r + kl[0] + 1)):
            Mdense[r, c] = dt[0, r, max_ku + c - r]
    u = 6 / Ly ** 2 * y * (Ly - y)               # bulk velocity 1
    T = np.ones(n)
    rho = np.ones(n)
    p = rho * T / gamma
    e = p / (gamma - 1) + Ma * Ma / 2 * rho * u * u
    vals = np.zeros(ch.shape, dtype=complex)
    vals[0, 0, 0] = e
    vals[1, 0, 0] = rho * u
    vals[4, 0, 0] = rho
    shape = y * (Ly - y)
    zabs = (ch.Nz - 1) // 2
    for iz in range(ch.dNz):
        wz = iz if iz < (ch.dNz + 1) // 2 else iz - ch.dNz
        if abs(wz) > zabs:
            continue
        for ix in range(ch.Xw):
            if ix > (ch.Nx - 1) // 2:
                continue
            kk = math.hypot(ix, wz)
            if kk == 0 or kk > kmax:
                continue
            for f in (1, 2, 3):
                c = (rng.standard_normal() + 1j * rng.standard_normal()) / math.sqrt(2)
                vals[f, iz, ix] = amp / kk ** 2 * c * shape
    # Hermitian symmetry on the kx = 0 plane keeps the fields real
    for iz in range(1, ch.dNz):
        jz = ch.dNz - iz
        if iz < jz:
            vals[:, jz, 0] = np.conj(vals[:, iz, 0])
    coeff = np.linalg.solve(Mdense, vals.reshape(-1, n).T).T.reshape(ch.shape)
    return coeff
