"""Python view of suzerain_amd::bsmbsm_solver (the C handle API).

    spec = "zgbsv" | "zcgbsvx,siter=3,diter=2,tolsc=..." | ...
    sol = BsmbsmSolver(spec, S=5, n=96, kl=6, ku=6, nrhs=1, nbatch=1152)
    model_assemble(... lu_ptr=sol.ptr("PAPT") ...)   # or pack into PAPT
    sol.supplied_PAPT()
    sol.supply_B(b)            # b: device tensor (nbatch, nrhs, N) complex128
    sol.solve("N")
    sol.demand_X(x)

mirrors suzerain::bsmbsm_solver (suzerain/bsmbsm_solver.hpp:70-330) for a
whole batch of systems at once.
"""
import ctypes as C

import numpy as np

from ._lib import load

_WHICH = {"LU": 0, "PAPT": 1, "PB": 2, "PX": 3, "ipiv": 4, "info": 5}
_STATS = {"siter": 0, "diter": 1, "tolsc": 2, "res": 3, "fact": 4, "apprx": 5}


def spec_normalize(spec):
    """Canonical string of a specification (raises ValueError like the
    reference's std::invalid_argument)."""
    buf = C.create_string_buffer(512)
    rc = load().sz_spec_normalize(spec.encode(), buf, 512)
    if rc:
        raise ValueError(buf.value.decode())
    return buf.value.decode()


def _hip():
    return C.CDLL("libamdhip64.so")


def device_to_host(ptr, nbytes, dtype):
    """Synchronous copy of raw device memory into a new numpy array."""
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
    h = _hip()
    h.hipDeviceSynchronize()
    rc = h.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                     C.c_size_t(nbytes), 2)
    if rc:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    return out


def host_to_device(ptr, arr):
    arr = np.ascontiguousarray(arr)
    h = _hip()
    rc = h.hipMemcpy(C.c_void_p(ptr), arr.ctypes.data_as(C.c_void_p),
                     C.c_size_t(arr.nbytes), 1)
    h.hipDeviceSynchronize()
    if rc:
        raise RuntimeError(f"hipMemcpy failed: {rc}")


def stats_of(handle, which, nbatch):
    """Per-system zcgbsvx outcome arrays of a solver handle (None when the
    solver is not zcgbsvx)."""
    p = load().sz_solver_zcgbsvx_stats(C.c_void_p(handle), _STATS[which])
    if not p:
        return None
    dtype = {"siter": np.int32, "diter": np.int32, "tolsc": np.float64,
             "res": np.float64, "fact": np.int8, "apprx": np.int32}[which]
    a = device_to_host(p, nbatch * np.dtype(dtype).itemsize, dtype)
    return a.view("S1") if which == "fact" else a


class BsmbsmSolver:
    def __init__(self, spec, S, n, kl, ku, nrhs=1, nbatch=1, stream=None):
        lib = load()
        self.geom = lib.sz_bsmbsm_construct(S, n, kl, ku).asdict()
        self.nrhs, self.nbatch = nrhs, nbatch
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        err = C.create_string_buffer(512)
        self._h = lib.sz_solver_create(spec.encode(), S, n, kl, ku, nrhs, nbatch,
                                       C.c_void_p(stream), err, 512)
        if not self._h:
            raise ValueError(err.value.decode())
        self.spec = spec_normalize(spec)

    def __getattr__(self, k):
        if k in ("S", "n", "kl", "ku", "ld", "N", "KL", "KU", "LD"):
            return self.geom[k]
        raise AttributeError(k)

    def storage(self, which):
        """(device pointer, leading dimension, per-system stride) of LU, PAPT,
        PB, PX, ipiv or info."""
        ld, stride = C.c_int(0), C.c_longlong(0)
        p = load().sz_solver_storage(self._h, _WHICH[which], C.byref(ld),
                                     C.byref(stride))
        return p, ld.value, stride.value

    def ptr(self, which):
        return self.storage(which)[0]

    def read(self, which):
        """Host copy of a storage block, shaped (nbatch, cols, ld)."""
        p, ld, stride = self.storage(which)
        dtype = np.int32 if which in ("ipiv", "info") else np.complex128
        a = device_to_host(p, stride * self.nbatch * np.dtype(dtype).itemsize,
                           dtype)
        return a.reshape(self.nbatch, stride)

    def write(self, which, arr):
        p, ld, stride = self.storage(which)
        host_to_device(p, arr)

    def stats(self, which):
        return stats_of(self._h, which, self.nbatch)

    def supplied_PAPT(self):
        load().sz_solver_supplied_PAPT(self._h)
        return self

    def apprx(self, acceptable):
        return bool(load().sz_solver_apprx(self._h, int(bool(acceptable))))

    def fact(self):
        return load().sz_solver_fact(self._h).decode()

    def supply_B(self, b):
        """b: device tensor (nbatch, nrhs, N) complex128."""
        rc = load().sz_solver_supply_B(self._h, C.c_void_p(b.data_ptr()),
                                       b.stride(1), 1, b.stride(0))
        if rc:
            raise RuntimeError(f"supply_B failed: {rc}")

    def demand_X(self, x):
        rc = load().sz_solver_demand_X(self._h, C.c_void_p(x.data_ptr()),
                                       x.stride(1), 1, x.stride(0))
        if rc:
            raise RuntimeError(f"demand_X failed: {rc}")

    def solve(self, trans, nrhs=None):
        """Returns 0 or the first nonzero per-system info; the message the
        reference would raise is kept in self.last_error."""
        err = C.create_string_buffer(512)
        rc = load().sz_solver_solve(self._h, trans.encode(),
                                    self.nrhs if nrhs is None else nrhs, err, 512)
        self.last_error = err.value.decode()
        return rc

    def first_info(self, clear=True):
        """First nonzero per-system info of any solve since the last clearing
        call (waits for the device), else 0."""
        return load().sz_solver_first_info(self._h, int(bool(clear)))

    def solve_async(self, trans, nrhs=None):
        return load().sz_solver_solve_async(self._h, trans.encode(),
                                            self.nrhs if nrhs is None else nrhs)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                load().sz_solver_free(self._h)
                self._h = None
        except Exception:
            pass
