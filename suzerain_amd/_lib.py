"""ctypes binding of libsuzerain_amd.so (the C ABI in include/suzerain_amd.h).

The library is the product; importing this module fails loudly when it is
missing instead of falling back to any CPU path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsuzerain_amd.so")

vp, i32, i64, f64 = C.c_void_p, C.c_int, C.c_longlong, C.c_double
pd = C.POINTER(C.c_double)
pi = C.POINTER(C.c_int)
ch = C.c_char


class sz_bsmbsm(C.Structure):
    """Mirror of the sz_bsmbsm / suzerain_bsmbsm geometry struct."""
    _fields_ = [(k, C.c_int) for k in
                ("S", "n", "kl", "ku", "ld", "N", "KL", "KU", "LD")]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class sz_imexop_scenario(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("Re", "Pr", "Ma", "alpha", "gamma")]


REF_NAMES = ("ux uy uz u2 uxux uxuy uxuz uyuy uyuz uzuz nu nuux nuuy nuuz nuu2 "
             "nuuxux nuuxuy nuuxuz nuuyuy nuuyuz nuuzuz ex_gradrho ey_gradrho "
             "ez_gradrho e_divm e_deltarho").split()


class sz_imexop_ref(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in REF_NAMES]


class sz_imexop_refld(C.Structure):
    _fields_ = [(k, C.c_int) for k in REF_NAMES]


class sz_isothermal_bc(C.Structure):
    _fields_ = [("enforce_lower", C.c_int), ("enforce_upper", C.c_int)] + [
        (k, C.c_double) for k in ("lower_T", "lower_u", "lower_v", "lower_w",
                                  "upper_T", "upper_u", "upper_v", "upper_w")]


# name: (restype, [argtypes])
SIGNATURES = {
    "sz_bsmbsm_construct": (sz_bsmbsm, [i32, i32, i32, i32]),
    "sz_bsmbsm_q": (i32, [i32, i32, i32]),
    "sz_bsmbsm_qinv": (i32, [i32, i32, i32]),
    "sz_bsmbsm_zaPxpby_batched": (i32, [ch, i32, i32, pd, vp, i32, i64, pd, vp,
                                        i32, i64, i32, vp]),
    "sz_bsmbsm_zpack_batched": (i32, [i32, i32, i32, i32, i32, i32, pd, vp, i32,
                                      i64, i32, i32, vp, i32, i64, i32, vp]),
    "sz_bsmbsm_zdpack_batched": (i32, [i32, i32, i32, i32, i32, i32, pd, vp,
                                       i32, i64, i32, i32, vp, i32, i64, i32,
                                       vp]),
    "sz_model_assemble_batched": (i32, [i32, i32, vp, i32, i32, i32, pi, pi,
                                        i64, vp, vp, vp, vp, f64, f64, i32, i32,
                                        vp, i32, i64, i32, vp]),
    "sz_zgbtrf_batched": (i32, [i32, i32, i32, vp, i32, i64, vp, i64, vp, i32,
                                vp]),
    "sz_cgbtrf_batched": (i32, [i32, i32, i32, vp, i32, i64, vp, i64, vp, i32,
                                vp]),
    "sz_zgbtrs_batched": (i32, [ch, i32, i32, i32, i32, vp, i32, i64, vp, i64,
                                vp, i32, i64, i32, vp]),
    "sz_cgbtrs_batched": (i32, [ch, i32, i32, i32, i32, vp, i32, i64, vp, i64,
                                vp, i32, i64, i32, vp]),
    "sz_zgbsv_batched": (i32, [ch, i32, i32, i32, i32, vp, i32, i64, vp, i64,
                               vp, i32, i64, vp, i32, vp]),
    "sz_htstretch1": (f64, [f64, f64, f64]),
    "sz_htstretch2": (f64, [f64, f64, f64]),
    "sz_bspline_breakpoints": (i32, [i32, i32, f64, f64, f64, pd]),
    "sz_bsplineop_alloc": (vp, [i32, i32, pd, i32, pi]),
    "sz_bsplineop_free": (None, [vp]),
    "sz_bsplineop_n": (i32, [vp]),
    "sz_bsplineop_k": (i32, [vp]),
    "sz_bsplineop_nderiv": (i32, [vp]),
    "sz_bsplineop_ld": (i32, [vp]),
    "sz_bsplineop_max_kl": (i32, [vp]),
    "sz_bsplineop_max_ku": (i32, [vp]),
    "sz_bsplineop_kl": (i32, [vp, i32]),
    "sz_bsplineop_ku": (i32, [vp, i32]),
    "sz_bsplineop_collocation_point": (f64, [vp, i32]),
    "sz_bsplineop_integration_coefficients": (i32, [vp, pd]),
    "sz_bsplineop_gram": (i32, [vp, pd]),
    "sz_mms_isothermal_channel": (None, [vp]),
    "sz_mms_evaluate": (i32, [vp, i32, vp, vp, vp, f64, vp, vp, vp]),
    "sz_bsplineop_D_T": (pd, [vp, i32]),
    "sz_bsplineop_D_T_device": (vp, [vp]),
    "sz_bsplineop_accumulate_complex_batched": (i32, [vp, i32, i32, pd, vp, i32,
                                                      i64, pd, vp, i32, i64,
                                                      vp]),
    "sz_bsplineop_accumulate_batched": (i32, [vp, i32, i32, f64, vp, i32, i64,
                                              f64, vp, i32, i64, vp]),
    "sz_diffwave_apply": (i32, [i32, i32, pd, vp, f64, f64, i32, i32, i32, i32,
                                i32, i32, i32, i32, i32, vp]),
    "sz_diffwave_accumulate": (i32, [i32, i32, pd, vp, pd, vp, f64, f64, i32,
                                     i32, i32, i32, i32, i32, i32, i32, i32,
                                     vp]),
    "sz_fft_plan_create": (vp, [i32, i32, i32, i32, i32, pi]),
    "sz_fft_plan_free": (None, [vp]),
    "sz_fft_mixed_elements": (i64, [vp]),
    "sz_fft_wave_to_physical": (i32, [vp, i32, vp, i64, vp, i64, vp, i64, vp]),
    "sz_fft_physical_to_wave": (i32, [vp, i32, vp, i64, vp, i64, vp, i64, vp]),
    "sz_zcgbsvx_workspace_bytes": (i64, [i32]),
    "sz_zcgbsvx_batched": (i32, [vp, vp, i32, ch, i32, i32, i32, vp, i32, i64,
                                 vp, vp, i32, i64, vp, i64, vp, i64, vp, i64,
                                 i32, i32, f64, vp, vp, vp, vp, i64, vp, vp,
                                 vp, i32, vp]),
    "sz_zcgbsvx_batched_noalloc": (i32, [vp, vp, i32, ch, i32, i32, i32, vp,
                                         i32, i64, vp, vp, i32, i64, vp, i64,
                                         vp, i64, vp, i64, i32, i32, f64, vp,
                                         vp, vp, vp, i64, vp, vp, i32, vp]),
    "sz_imexop_create": (vp, [vp, vp, pi]),
    "sz_imexop_free": (None, [vp]),
    "sz_imexop_nterms": (i32, [vp]),
    "sz_rholut_imexop_packc_batched": (i32, [vp, pd, vp, vp, vp, vp, i32, i32,
                                             i32, i32, i32, vp, vp, i32, i64,
                                             i32, vp, vp]),
    "sz_rholut_imexop_packf_batched": (i32, [vp, pd, vp, vp, vp, vp, i32, i32,
                                             i32, i32, i32, vp, vp, i32, i64,
                                             i32, vp, vp]),
    "sz_rholut_imexop_accumulate_batched": (i32, [vp, pd, vp, vp, vp, vp, vp,
                                                  vp, i64, pd, vp, vp, i64,
                                                  i32, vp]),
    "sz_isothermal_rhs_batched": (i32, [i32, i32, i32, i32, i32, i32, i32, i32,
                                        vp, i64, i32, vp]),
    "sz_ns_workspace_bytes": (i64, [i32, i32, i32]),
    "sz_ns_references": (i32, [vp, i32, i32, i32, vp, i64, vp, vp, vp]),
    "sz_ns_rhs": (i32, [vp, i32, i32, i32, i32, i32, vp, i64, vp, i64, vp,
                        i64, vp, vp, vp, vp, vp, vp, vp]),
    "sz_spec_normalize": (i32, [C.c_char_p, C.c_char_p, i32]),
    "sz_solver_create": (vp, [C.c_char_p, i32, i32, i32, i32, i32, i32, vp,
                              C.c_char_p, i32]),
    "sz_solver_free": (None, [vp]),
    "sz_solver_storage": (vp, [vp, i32, pi, C.POINTER(C.c_longlong)]),
    "sz_solver_first_info": (i32, [vp, i32]),
    "sz_solver_zcgbsvx_stats": (vp, [vp, i32]),
    "sz_solver_supply_B": (i32, [vp, vp, i32, i32, i64]),
    "sz_solver_demand_X": (i32, [vp, vp, i32, i32, i64]),
    "sz_solver_supplied_PAPT": (None, [vp]),
    "sz_solver_apprx": (i32, [vp, i32]),
    "sz_solver_fact": (ch, [vp]),
    "sz_solver_solve": (i32, [vp, ch, i32, C.c_char_p, i32]),
    "sz_solver_solve_async": (i32, [vp, ch, i32]),
}

_lib = None


def load():
    """Return the loaded library, declaring every signature once."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(
            f"{LIBPATH} is missing: build it with `python -m suzerain_amd.build`"
            " (there is no CPU fallback)")
    # PyTorch (device memory and streams) brings its own HIP runtime under
    # the same soname: load it first, so the library binds to that one
    # runtime instead of loading a second copy from /opt/rocm next to it
    import torch  # noqa: F401
    lib = C.CDLL(LIBPATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def cplx(z):
    """Host complex scalar as the (re, im) double pair the ABI expects."""
    z = complex(z)
    return (C.c_double * 2)(z.real, z.imag)


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")
    return rc
