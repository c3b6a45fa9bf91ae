"""GPU parity of the dealiased wave <-> physical transforms against numpy's
pocketfft (FFTW sign and normalization conventions: unnormalized, +i
backward), plus the size-independent round trip
physical_to_wave(wave_to_physical(W)) = dNx dNz W on dealiased W."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dealiased_wave(rng, nf, Ny, Nx, dNx, Nz, dNz):
    Xw = dNx // 2 + 1
    W = rng.standard_normal((nf, dNz, Xw, Ny)) + 1j * rng.standard_normal((nf, dNz, Xw, Ny))
    kx = np.arange(Xw)
    kz = np.arange(dNz)
    wz = np.where(kz < dNz // 2 + 1, kz, dNz - kz)
    keep = (kx[None, :] <= (Nx - 1) // 2) & (wz[:, None] <= (Nz - 1) // 2)
    W *= keep[None, :, :, None]
    return W, keep


def backward_numpy(W, dNx, dNz):
    # W (nf, dNz, Xw, Ny): z c2c (+), then x c2r (+), unnormalized
    T = np.fft.ifft(W, axis=1) * dNz
    P = np.fft.irfft(T, n=dNx, axis=2) * dNx          # (nf, dNz, dNx, Ny)
    return np.transpose(P, (0, 3, 1, 2))              # (nf, Ny, dNz, dNx)


def forward_numpy(P, keep):
    # P (nf, Ny, dNz, dNx) -> r2c along x, c2c along z, dealias
    X = np.fft.rfft(P, axis=3)                        # (nf, Ny, dNz, Xw)
    Z = np.fft.fft(X, axis=2)
    W = np.transpose(Z, (0, 2, 3, 1))                 # (nf, dNz, Xw, Ny)
    return W * keep[None, :, :, None]


@pytest.mark.parametrize("Ny,Nx,Nz", [(96, 96, 96), (12, 32, 48), (7, 64, 128),
                                      (7, 192, 192), (6, 144, 144), (5, 384, 384),
                                      (9, 4, 6)])
def test_transforms_match_numpy(gpu, Ny, Nx, Nz):
    _check_transforms(gpu, Ny, Nx, Nz)


@pytest.mark.parametrize("Ny,Nx,Nz", [(96, 96, 96), (7, 192, 192)])
def test_eight_line_z_tiles(gpu, monkeypatch, Ny, Nx, Nz):
    # SZ_FFT_ZLINES=8 selects the 8-line z tile at plan creation
    monkeypatch.setenv("SZ_FFT_ZLINES", "8")
    _check_transforms(gpu, Ny, Nx, Nz)


def _check_transforms(gpu, Ny, Nx, Nz):
    import torch
    from suzerain_amd.fft import PencilFFT
    print(f"case Ny={Ny} Nx={Nx} Nz={Nz}", flush=True)
    dNx, dNz = 3 * Nx // 2, 3 * Nz // 2
    rng = np.random.default_rng(Ny + Nx)
    nf = 3
    W, keep = dealiased_wave(rng, nf, Ny, Nx, dNx, Nz, dNz)
    fft = PencilFFT(Ny, Nx, dNx, Nz, dNz)
    Wd = torch.from_numpy(W).to(gpu)
    Pd = torch.empty((nf, Ny, dNz, dNx), dtype=torch.float64, device=gpu)
    fft.wave_to_physical(Wd, Pd)
    P = Pd.cpu().numpy()
    Pref = backward_numpy(W, dNx, dNz)
    scale = np.abs(Pref).max()
    np.testing.assert_allclose(P, Pref, rtol=0, atol=1e-14 * scale)
    # forward of an arbitrary physical field
    Pr = rng.standard_normal((nf, Ny, dNz, dNx))
    Wout = torch.full((nf, dNz, dNx // 2 + 1, Ny), np.nan + 0j,
                      dtype=torch.complex128, device=gpu)
    fft.physical_to_wave(torch.from_numpy(Pr).to(gpu), Wout)
    Wref = forward_numpy(Pr, keep)
    np.testing.assert_allclose(Wout.cpu().numpy(), Wref, rtol=0,
                               atol=1e-13 * np.abs(Wref).max())
    # round trip on a dealiased spectrum of a real field (the kx = 0 plane
    # of an arbitrary W is not Hermitian, so only such spectra survive)
    Wh = np.ascontiguousarray(Wref / (dNx * dNz))
    Pd2 = torch.empty_like(Pd)
    fft.wave_to_physical(torch.from_numpy(Wh).to(gpu), Pd2)
    W2 = torch.zeros_like(Wd)
    fft.physical_to_wave(Pd2, W2)
    np.testing.assert_allclose(W2.cpu().numpy() / (dNx * dNz), Wh, rtol=0,
                               atol=1e-13 * np.abs(Wh).max())
