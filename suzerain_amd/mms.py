"""Manufactured solutions (the nsctpl_rholut fields of apps/perfect): Python
view of the sz_mms parameter struct and its device evaluation.

    m = Manufactured.isothermal_channel()      # nsctpl_rholut.hpp:991-1104
    m.set(alpha=1.0)                           # scenario overrides
    ch.set_manufactured(m)                     # forcing in the channel's N
    ch.manufactured_state(0.0)                 # exact state at t = 0
"""
import ctypes as C

from ._lib import load, check

TERMS = ("t0", "x", "y", "z", "xy", "xz", "yz")
PRIMITIVES = ("rho", "u", "v", "w", "T")
SCALARS = ("alpha", "beta", "gamma", "Ma", "Pr", "Re", "Lx", "Ly", "Lz")


class sz_mms_term(C.Structure):
    _fields_ = [(k, C.c_double) for k in "abcdefg"]


class sz_mms_primitive(C.Structure):
    _fields_ = [(t, sz_mms_term) for t in TERMS]


class sz_mms(C.Structure):
    _fields_ = ([(k, C.c_double) for k in SCALARS] +
                [(p, sz_mms_primitive) for p in PRIMITIVES])


class Manufactured:
    """A manufactured solution's parameters (sz_mms)."""

    def __init__(self, struct=None):
        self.s = struct if struct is not None else sz_mms()

    @classmethod
    def isothermal_channel(cls):
        s = sz_mms()
        lib = load()
        lib.sz_mms_isothermal_channel.argtypes = [C.c_void_p]
        lib.sz_mms_isothermal_channel.restype = None
        lib.sz_mms_isothermal_channel(C.byref(s))
        return cls(s)

    def set(self, **kw):
        """Scenario and domain values (alpha, beta, gamma, Ma, Pr, Re, Lx,
        Ly, Lz) or term parameters named like the reference's options,
        e.g. rho_a_xy=0 (term "0" is t0)."""
        for k, v in kw.items():
            if k in SCALARS:
                setattr(self.s, k, float(v))
                continue
            field, param, term = k.split("_", 2)
            term = "t0" if term == "0" else term
            setattr(getattr(getattr(self.s, field), term), param, float(v))
        return self

    def parameters(self):
        """Flat dict: scalars and <field>_<param>_<term> entries."""
        out = {k: getattr(self.s, k) for k in SCALARS}
        for p in PRIMITIVES:
            for t in TERMS:
                term = getattr(getattr(self.s, p), t)
                for q in "abcdefg":
                    out[f"{p}_{q}_{'0' if t == 't0' else t}"] = getattr(term, q)
        return out

    def evaluate(self, x, y, z, t, stream=None):
        """(conserved, Q) at device float64 points: two (5, n) tensors."""
        import torch
        x, y, z = (a.to(torch.float64).contiguous() for a in (x, y, z))
        n = x.numel()
        cons = torch.empty((5, n), dtype=torch.float64, device=x.device)
        Q = torch.empty_like(cons)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        lib = load()
        f = lib.sz_mms_evaluate
        f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                      C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        f.restype = C.c_int
        check(f(C.byref(self.s), n, x.data_ptr(), y.data_ptr(), z.data_ptr(), t,
                cons.data_ptr(), Q.data_ptr(), C.c_void_p(stream)), "sz_mms_evaluate")
        return cons, Q
