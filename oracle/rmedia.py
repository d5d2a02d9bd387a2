"""CPU restatement of the reference's random-media synthesis -- TEST
INFRASTRUCTURE ONLY (the checker of the HIP generator and the CPU baseline).

src/shared/m_rmedia.f90:40-173 (rmedia__3dgen_ani; rmedia__3dgen is the
isotropic case ax = ay = az = a):
  1. uniform noise media(i,j,k) in [0, 1);
  2. forward FFT, b(k) = dV * sum a(x) exp(-i k.x) (m_fk.f90 x2k3d);
  3. whitening b = b / |b|;
  4. filtering b = b * sqrt(psdf(kx, ky, kz) * N * dV);
  5. b(0) = 0, inverse FFT, real part / (N * dV) (m_fk.f90 k2x3d);
  6. mean removed, scaled so that the rms equals epsil.
The PSDFs are m_rmedia.f90:590-732 (psdf_*3d and psdf_ani_*3d).  The noise
comes from the Fortran runtime's random_number, whose stream is not part of
the algorithm: realisations are compared through their spectra.  After
steps 3-6 every non-zero wavenumber carries |F(k)|^2 = C psdf(k) with one
constant C, so spectral_ratio() below is flat for any seed.

Parity status: psdf_* and wavenumbers() are checked against the reference's
own values (tests/golden/rmedia_golden.txt, tests/test_oracle_rmedia.py);
generate() against the flat-spectrum property of the reference's own two
realisations in the same file.
"""
from __future__ import annotations

import math

import numpy as np

PI = math.pi
GAUSS, EXPONENTIAL, VON_KARMAN = 1, 2, 3


def _gamma(x):
    return math.gamma(float(np.float32(x)))


def psdf_ani(ptype, mx, my, mz, ax, ay, az, eps, kappa=0.5):
    """rmedia__psdf_ani_{gauss,exp,vonKarman}3d (m_rmedia.f90:682-732),
    float64 arithmetic on float32 inputs; works on arrays."""
    f = np.float64
    mx, my, mz = (np.asarray(v, np.float32).astype(f) for v in (mx, my, mz))
    ax, ay, az, eps, kappa = (f(np.float32(v)) for v in (ax, ay, az, eps, kappa))
    q = (mx * ax) ** 2 + (my * ay) ** 2 + (mz * az) ** 2
    if ptype == GAUSS:
        return eps * eps * ax * ay * az * PI ** 1.5 * np.exp(-q * 0.25)
    if ptype == EXPONENTIAL:
        wk = 1.0 + q
        return 8 * eps * eps * ax * ay * az * PI / (wk * wk)
    return (8 * PI ** 1.5 * eps * eps * ax * ay * az * _gamma(kappa + 1.5)) / \
        (_gamma(kappa) * (1.0 + q) ** (kappa + 1.5))


def psdf_iso(ptype, m, a, eps, kappa=0.5):
    """rmedia__psdf_{gauss,exp,vonKarman}3d (m_rmedia.f90:590-679)"""
    return psdf_ani(ptype, m, 0.0, 0.0, a, a, a, eps, kappa)


def wavenumbers(n, d):
    """kx of fk__x2k_3d (m_fk.f90 x2k3d_4): 2 pi (i-1)/(n d) up to n/2+1,
    then -2 pi (n-i+1)/(n d); single precision"""
    nd = np.float32(np.float32(n) * np.float32(d))
    i = np.arange(n)
    v = np.where(i <= n // 2, 2 * PI * i, -2 * PI * (n - i)) / np.float64(nd)
    return v.astype(np.float32)


def generate(ptype, shape, d, a, eps, kappa, seed):
    """A realisation on a grid of shape (n1, n2, n3) (axis 0 fastest in the
    reference's media(nx, ny, nz); here a numpy array indexed [i1, i2, i3]),
    spacings d = (d1, d2, d3), correlation lengths a = (a1, a2, a3).
    Float64 FFTs; returns float32."""
    n1, n2, n3 = shape
    rng = np.random.default_rng(seed)
    noise = rng.random(shape)
    dv = float(np.float32(d[0])) * float(np.float32(d[1])) * float(np.float32(d[2]))
    b = np.fft.fftn(noise) * dv
    with np.errstate(divide="ignore", invalid="ignore"):
        b = b / np.abs(b)
    k1, k2, k3 = (wavenumbers(n, dd) for n, dd in zip(shape, d))
    p = psdf_ani(ptype, k1[:, None, None], k2[None, :, None], k3[None, None, :], *a, eps, kappa)
    b = b * np.sqrt(p * n1 * n2 * n3 * dv)
    b[0, 0, 0] = 0.0
    m = np.real(np.fft.ifftn(b)) * (n1 * n2 * n3) / (n1 * n2 * n3 * dv)
    m = m - m.mean()
    m = m * (float(np.float32(eps)) / math.sqrt(float(np.mean(m * m))))
    return m.astype(np.float32)


def spectral_ratio(field, ptype, d, a, eps, kappa, floor=1e-8):
    """|F(k)|^2 / psdf(k) over the non-zero wavenumbers whose PSDF is above
    floor * max: (ratio / median ratio) and the mask used"""
    f = np.asarray(field, np.float64)
    F2 = np.abs(np.fft.fftn(f)) ** 2
    k1, k2, k3 = (wavenumbers(n, dd) for n, dd in zip(f.shape, d))
    p = psdf_ani(ptype, k1[:, None, None], k2[None, :, None], k3[None, None, :], *a, eps, kappa)
    mask = p > floor * p.max()
    mask[0, 0, 0] = False
    r = F2[mask] / p[mask]
    return r / np.median(r), mask
