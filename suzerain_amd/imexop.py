"""Compressible implicit operator M + phi L on the device.

Python view of the sz_imexop C handle: the batched counterparts of
suzerain_rholut_imexop_packc/packf/accumulate (suzerain/rholut_imexop.h:
186-441) and of IsothermalPATPTEnforcer (apps/perfect/
operator_hybrid_isothermal.cpp:396-526).  Reference profiles are a float64
device tensor shaped (26, n) in suzerain_rholut_imexop_ref member order.
"""
import ctypes as C

from ._lib import (load, cplx, check, sz_bsmbsm, sz_imexop_scenario,
                   sz_imexop_ref, sz_imexop_refld, sz_isothermal_bc, REF_NAMES)

__all__ = ["Imexop", "REF_NAMES", "isothermal_bc"]


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def isothermal_bc(lower=(1.0, 0.0, 0.0, 0.0), upper=(1.0, 0.0, 0.0, 0.0),
                  enforce_lower=True, enforce_upper=True):
    """Wall (T, u, v, w) at the lower and upper walls."""
    return sz_isothermal_bc(int(enforce_lower), int(enforce_upper), *lower,
                            *upper)


class Imexop:
    def __init__(self, op, Re, Pr, Ma, alpha, gamma):
        """op: suzerain_amd.BsplineOp with nderiv >= 2 (kept alive here)."""
        self.op = op
        self.scenario = (Re, Pr, Ma, alpha, gamma)
        sc = sz_imexop_scenario(Re, Pr, Ma, alpha, gamma)
        info = C.c_int(0)
        self._h = load().sz_imexop_create(C.byref(sc), op._w, C.byref(info))
        if not self._h:
            raise ValueError(f"sz_imexop_create failed: {info.value}")
        self.nterms = load().sz_imexop_nterms(self._h)

    @staticmethod
    def _refs(refs):
        assert refs.dtype.is_floating_point and refs.shape[0] == 26
        r = sz_imexop_ref(*[refs[i].data_ptr() for i in range(26)])
        ld = sz_imexop_refld(*([refs.stride(1)] * 26))
        return r, ld

    def pack(self, phi, km, kn, refs, papt, ldpapt, stride, nsys,
             positions=(0, 1, 2, 3, 4), bc=None, packf=False):
        """Write P (M + phi L)^T P^T for nsys systems into raw band storage at
        device address papt (packc) or LU storage (packf: band at +KL).
        Returns the BSMBSM geometry as a dict."""
        r, ld = self._refs(refs)
        g = sz_bsmbsm()
        f = (load().sz_rholut_imexop_packf_batched if packf
             else load().sz_rholut_imexop_packc_batched)
        rc = f(self._h, cplx(phi), C.c_void_p(km.data_ptr()),
               C.c_void_p(kn.data_ptr()), C.byref(r), C.byref(ld), *positions,
               C.byref(g), C.c_void_p(papt), ldpapt, stride, nsys,
               C.byref(bc) if bc is not None else None, _stream())
        check(rc, "sz_rholut_imexop_pack")
        return g.asdict()

    def accumulate(self, phi, km, kn, refs, x, beta, y, nsys=None,
                   inoff=None, outoff=None):
        """y_f <- (M + phi L) x + beta y_f for complex128 tensors x, y shaped
        (5, nsys, n) (field, system, y) or sequences of 5 per-field tensors
        (None omits a field) with optional per-system offsets."""
        r, ld = self._refs(refs)
        n = self.op.n
        for t in (list(x) if not hasattr(x, "dim") else [x]) + \
                (list(y) if not hasattr(y, "dim") else [y]):
            if t is not None and t.stride(-1) != 1:
                raise ValueError("pencils must be contiguous along y")
        if hasattr(x, "dim"):
            nsys = x.shape[1] if nsys is None else nsys
            xs = [x[f] for f in range(5)]
            ys = [y[f] for f in range(5)]
            sx, sy = x.stride(1), y.stride(1)
        else:
            xs, ys = list(x), list(y)
            sx = sy = n
        ins = (C.c_void_p * 5)(*[t.data_ptr() if t is not None else None
                                 for t in xs])
        outs = (C.c_void_p * 5)(*[t.data_ptr() if t is not None else None
                                  for t in ys])
        rc = load().sz_rholut_imexop_accumulate_batched(
            self._h, cplx(phi), C.c_void_p(km.data_ptr()),
            C.c_void_p(kn.data_ptr()), C.byref(r), C.byref(ld), ins,
            C.c_void_p(inoff.data_ptr()) if inoff is not None else None, sx,
            cplx(beta), outs,
            C.c_void_p(outoff.data_ptr()) if outoff is not None else None, sy,
            nsys, _stream())
        check(rc, "sz_rholut_imexop_accumulate_batched")
        return y

    @staticmethod
    def isothermal_rhs(S, n, pb, stride, nsys, positions=(0, 1, 2, 3),
                       enforce_lower=True, enforce_upper=True):
        rc = load().sz_isothermal_rhs_batched(
            S, n, *positions, int(enforce_lower), int(enforce_upper),
            C.c_void_p(pb), stride, nsys, _stream())
        check(rc, "sz_isothermal_rhs_batched")

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                load().sz_imexop_free(self._h)
                self._h = None
        except Exception:
            pass
