"""numpy helpers for the tests: layout conversion and independent spectral math.

Everything here is written from the mathematics (numpy.fft, curl, divergence),
not from the reference's loops, so that checks against it pin the oracle.
"""
import numpy as np


def pad_real(real):
    nx, ny, nz = real.shape
    a = np.zeros((nx, ny, nz + 2))
    a[:, :, :nz] = real
    return a


def to_complex(padded):
    """padded k-space field (NX, NY, NZ+2) -> complex (NX, NY, NZ/2+1)."""
    return np.ascontiguousarray(padded).view(np.complex128)


def from_complex(c):
    return np.ascontiguousarray(c).view(np.float64).copy()


def wavevectors(shape, dx=1.0):
    """Signed integer wavenumbers and physical k arrays of the half spectrum."""
    nx, ny, nz = shape
    ix = np.fft.fftfreq(nx, 1.0 / nx).astype(int)
    iy = np.fft.fftfreq(ny, 1.0 / ny).astype(int)
    iz = np.arange(nz // 2 + 1)
    # the reference keeps +N/2 for the Nyquist index (Calc_KX_Ooura: i > N/2 -> i - N)
    ix[nx // 2] = nx // 2
    iy[ny // 2] = ny // 2
    KX, KY, KZ = np.meshgrid(ix, iy, iz, indexing="ij")
    wx, wy, wz = 2 * np.pi / (nx * dx), 2 * np.pi / (ny * dx), 2 * np.pi / (nz * dx)
    return (KX, KY, KZ), (wx * KX, wy * KY, wz * KZ)


def kept_mask(shape):
    """2/3 rule: a mode survives iff |kx| < TRN_X, |ky| < TRN_Y, kz < TRN_Z, TRN = (N+2)/3."""
    (KX, KY, KZ), _ = wavevectors(shape)
    trn = [(n + 2) // 3 for n in shape]
    return (np.abs(KX) < trn[0]) & (np.abs(KY) < trn[1]) & (KZ < trn[2])


def band_mask(shape, kmax):
    (KX, KY, KZ), _ = wavevectors(shape)
    return (np.abs(KX) <= kmax) & (np.abs(KY) <= kmax) & (KZ <= kmax)


def random_solenoidal(shape, rng, kmax=None, amp=1.0, dc=(0.0, 0.0, 0.0)):
    """Random real solenoidal velocity (3, NX, NY, NZ), band-limited to |k_i| <= kmax."""
    nx, ny, nz = shape
    u = rng.standard_normal((3,) + shape) * amp
    uk = np.fft.rfftn(u, axes=(1, 2, 3))
    _, (kx, ky, kz) = wavevectors(shape)
    k2 = kx * kx + ky * ky + kz * kz
    ik2 = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    dot = (kx * uk[0] + ky * uk[1] + kz * uk[2]) * ik2
    uk = uk - np.stack([kx, ky, kz]) * dot
    mask = kept_mask(shape) if kmax is None else band_mask(shape, kmax)
    # drop the Nyquist planes entirely so the field is exactly Hermitian-representable
    mask = mask & (np.abs(wavevectors(shape)[0][0]) < nx // 2) & (np.abs(wavevectors(shape)[0][1]) < ny // 2) \
        & (wavevectors(shape)[0][2] < nz // 2)
    uk = uk * mask
    for d in range(3):
        uk[d, 0, 0, 0] = dc[d] * nx * ny * nz
    u = np.fft.irfftn(uk, s=shape, axes=(1, 2, 3))
    return u


def spectral_grad(f, dx=1.0):
    """Spectral gradient of a real periodic field (band-limited below Nyquist)."""
    shape = f.shape
    fk = np.fft.rfftn(f)
    _, (kx, ky, kz) = wavevectors(shape, dx)
    return [np.fft.irfftn(1j * k * fk, s=shape, axes=(0, 1, 2)) for k in (kx, ky, kz)]


def curl_k(uk, shape, dx=1.0):
    """i k x u_k."""
    _, (kx, ky, kz) = wavevectors(shape, dx)
    return np.stack([1j * (ky * uk[2] - kz * uk[1]), 1j * (kz * uk[0] - kx * uk[2]), 1j * (kx * uk[1] - ky * uk[0])])
