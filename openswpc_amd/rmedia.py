"""Seeded von Karman random media by spectral filtering.

Same construction as the reference's generator (src/shared/m_rmedia.f90:40-173:
random phases shaped by the power spectral density function, transformed to
space, normalised to the requested standard deviation), with numpy's
PCG64 stream and scipy's FFT.  The reference draws its noise from the
compiler's random_number, so the fields are not reproducible from it; what is
checked is the statistics (standard deviation and the spectral decay).

PSDF of a 3D von Karman medium, correlation length a, Hurst number kappa:
    P(k) = 8 pi^(3/2) eps^2 a^3 Gamma(kappa+3/2) / (Gamma(kappa) (1 + k^2 a^2)^(kappa+3/2))
Only the shape (1 + k^2 a^2)^-(kappa+3/2) matters after normalisation.
"""
from __future__ import annotations

import numpy as np


def von_karman(shape_jik, dx, dy, dz, a, eps, kappa, seed, workers=None):
    """Random field of shape (nj, ni, nk) = numpy view of a (k,i,j) array,
    zero mean and standard deviation eps, float32."""
    import scipy.fft as sfft

    nj, ni, nk = shape_jik
    rng = np.random.default_rng(seed)
    kj = 2 * np.pi * np.fft.fftfreq(nj, d=dy).astype(np.float32)
    ki = 2 * np.pi * np.fft.fftfreq(ni, d=dx).astype(np.float32)
    kk = 2 * np.pi * np.fft.rfftfreq(nk, d=dz).astype(np.float32)
    expo = np.float32(-(kappa + 1.5) / 2.0)  # sqrt of the PSDF
    a2 = np.float32(a * a)
    spec = np.empty((nj, ni, kk.size), np.complex64)
    for j in range(nj):  # plane by plane keeps the temporaries small
        k2 = ki[:, None] ** 2 + kj[j] ** 2 + kk[None, :] ** 2
        amp = (1 + k2 * a2) ** expo
        re = rng.standard_normal(amp.shape, dtype=np.float32)
        im = rng.standard_normal(amp.shape, dtype=np.float32)
        spec[j] = (re + 1j * im) * amp
    spec[0, 0, 0] = 0
    field = sfft.irfftn(spec, s=(nj, ni, nk), axes=(0, 1, 2), workers=workers or -1)
    del spec
    field = field.astype(np.float32, copy=False)
    field -= np.float32(field.mean(dtype=np.float64))
    std = float(field.std(dtype=np.float64))
    if std > 0:
        field *= np.float32(eps / std)
    return field
