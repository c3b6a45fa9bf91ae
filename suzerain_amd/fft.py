"""Dealiased wave <-> physical transforms (the pencil_grid call surface).

    fft = PencilFFT(Ny, Nx, dNx, Nz, dNz)
    fft.wave_to_physical(wave, phys)     # wave (nf, dNz, dNx/2+1, Ny) c128
    fft.physical_to_wave(phys, wave)     # phys (nf, Ny, dNz, dNx) f64

Tensor shapes are numpy C order, i.e. the reference's column-major (Y, X, Z)
wave and (X, Z, Y) physical layouts (suzerain/pencil_grid.hpp:200-237).
"""
import ctypes as C

from ._lib import load, check


class PencilFFT:
    def __init__(self, Ny, Nx, dNx, Nz, dNz):
        info = C.c_int(0)
        self._p = load().sz_fft_plan_create(Ny, Nx, dNx, Nz, dNz, C.byref(info))
        if not self._p:
            raise ValueError(f"sz_fft_plan_create failed: {info.value}")
        self.Ny, self.Nx, self.dNx, self.Nz, self.dNz = Ny, Nx, dNx, Nz, dNz
        self.mixed_elems = load().sz_fft_mixed_elements(self._p)
        self._mixed = None

    def chi(self):
        return 1.0 / (self.dNx * self.dNz)

    def _work(self, nf, device):
        import torch
        if self._mixed is None or self._mixed.shape[0] < nf:
            self._mixed = torch.empty((nf, self.mixed_elems),
                                      dtype=torch.complex128, device=device)
        return self._mixed

    def _stream(self):
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _check(self, wave, phys):
        """Every field must be one dense C-ordered block: the kernels address
        wave (dNz, dNx/2+1, Ny) and physical (Ny, dNz, dNx) fields from the
        field base pointer with those extents."""
        wshape = (self.dNz, self.dNx // 2 + 1, self.Ny)
        pshape = (self.Ny, self.dNz, self.dNx)
        if tuple(wave.shape[1:]) != wshape or not wave[0].is_contiguous():
            raise ValueError(f"wave fields must be contiguous {wshape}")
        if tuple(phys.shape[1:]) != pshape or not phys[0].is_contiguous():
            raise ValueError(f"physical fields must be contiguous {pshape}")
        if wave.shape[0] != phys.shape[0]:
            raise ValueError("field counts differ")

    def wave_to_physical(self, wave, phys):
        self._check(wave, phys)
        nf = wave.shape[0]
        mix = self._work(nf, wave.device)
        rc = load().sz_fft_wave_to_physical(
            self._p, nf, C.c_void_p(wave.data_ptr()), wave.stride(0),
            C.c_void_p(mix.data_ptr()), mix.stride(0), C.c_void_p(phys.data_ptr()),
            phys.stride(0), self._stream())
        return check(rc, "sz_fft_wave_to_physical")

    def physical_to_wave(self, phys, wave):
        self._check(wave, phys)
        nf = phys.shape[0]
        mix = self._work(nf, phys.device)
        rc = load().sz_fft_physical_to_wave(
            self._p, nf, C.c_void_p(phys.data_ptr()), phys.stride(0),
            C.c_void_p(mix.data_ptr()), mix.stride(0), C.c_void_p(wave.data_ptr()),
            wave.stride(0), self._stream())
        return check(rc, "sz_fft_physical_to_wave")

    def __del__(self):
        try:
            if getattr(self, "_p", None):
                load().sz_fft_plan_free(self._p)
                self._p = None
        except Exception:
            pass
