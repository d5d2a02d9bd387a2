"""The random-media oracle (oracle/rmedia.py) against the reference's own
m_rmedia.f90/m_fk.f90 (tests/golden/rmedia_golden.txt, made by
tests/golden/make_rmedia_golden.py): PSDF values, FFT wavenumbers, and the
flat-spectrum property |F(k)|^2 = C psdf(k) of the reference's own two
realisations, which is what the HIP generator is held to."""
import os

import numpy as np
import pytest

from oracle import rmedia as R

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmedia_golden.txt")


def _f(h):
    return np.array([int(h, 16)], np.uint32).view(np.float32)[0]


def _load():
    iso, ani, ks, fields = [], [], {}, []
    lines = [l for l in open(GOLD) if not l.startswith("#")]
    it = iter(lines)
    for line in it:
        w = line.split()
        if not w:
            continue
        if w[0] == "iso":
            iso.append((int(w[1]), *[_f(h) for h in w[2:]]))
        elif w[0] == "ani":
            ani.append((int(w[1]), *[_f(h) for h in w[2:]]))
        elif w[0] == "k":
            ks[(int(w[1]), _f(w[2]))] = np.array([_f(h) for h in w[3:]], np.float32)
        elif w[0] == "field":
            kind, ptype = w[1], int(w[2])
            n = tuple(int(v) for v in w[3:6])
            d = tuple(float(v) for v in w[6:9])
            a = tuple(float(v) for v in w[9:12])
            eps, kappa = float(w[12]), float(w[13])
            vals = []
            while len(vals) < n[0] * n[1] * n[2]:
                vals += [_f(h) for h in next(it).split()]
            media = np.array(vals, np.float32).reshape(n[::-1]).transpose(2, 1, 0)   # media(nx, ny, nz)
            fields.append(dict(kind=kind, ptype=ptype, d=d, a=a, eps=eps, kappa=kappa, media=media))
    return iso, ani, ks, fields


@pytest.fixture(scope="module")
def gold():
    return _load()


def test_isotropic_psdf(gold):
    iso, _, _, _ = gold
    assert len(iso) == 6 * 3 * 5
    for row in iso:
        ptype = row[0]
        if ptype == 3:
            _, m, a, kappa, want = row
        else:
            _, m, a, want = row
            kappa = 0.5
        got = np.float32(R.psdf_iso(ptype, m, a, 0.05, kappa))   # the reference underflows in single
        assert abs(got - want) <= 3e-6 * abs(want) + 1e-37, (row, got)


def test_anisotropic_psdf(gold):
    _, ani, _, _ = gold
    assert len(ani) == 18
    for ptype, mx, my, mz, want in ani:
        got = np.float32(R.psdf_ani(ptype, mx, my, mz, 2.0, 1.0, 0.5, 0.07, 0.3))
        assert abs(got - want) <= 3e-6 * abs(want) + 1e-37, (ptype, mx, got, want)


def test_wavenumbers(gold):
    _, _, ks, _ = gold
    assert len(ks) == 4
    for (n, d), want in ks.items():
        assert np.array_equal(R.wavenumbers(n, d), want), (n, d)


def test_reference_realisations_have_flat_spectra(gold):
    """the property the GPU generator is checked by holds for the
    reference's own output: mean 0, rms eps, |F|^2 proportional to the PSDF"""
    _, _, _, fields = gold
    assert len(fields) == 2
    for fd in fields:
        m = fd["media"].astype(np.float64)
        assert abs(m.mean()) < 1e-6 * fd["eps"]
        assert abs(np.sqrt(np.mean(m * m)) / fd["eps"] - 1) < 1e-5
        r, mask = R.spectral_ratio(fd["media"], fd["ptype"], fd["d"], fd["a"], fd["eps"], fd["kappa"], 1e-6)
        assert mask.sum() > 0.5 * m.size
        assert np.max(np.abs(r - 1)) < 1e-3, (fd["kind"], np.max(np.abs(r - 1)))


def test_flatness_detects_a_wrong_psdf(gold):
    """control: the same reference field against a PSDF with another
    correlation length or Hurst number is far from flat"""
    _, _, _, fields = gold
    fd = fields[0]
    a2 = tuple(2 * v for v in fd["a"])
    r, _ = R.spectral_ratio(fd["media"], fd["ptype"], fd["d"], a2, fd["eps"], fd["kappa"], 1e-6)
    assert np.max(np.abs(r - 1)) > 0.3
    r, _ = R.spectral_ratio(fd["media"], fd["ptype"], fd["d"], fd["a"], fd["eps"], 1.0, 1e-6)
    assert np.max(np.abs(r - 1)) > 0.3


@pytest.mark.parametrize("ptype", [1, 2, 3])
def test_oracle_generator_has_the_same_properties(ptype):
    d, a = (0.5, 0.4, 0.3), (2.0, 1.5, 1.0)
    m = R.generate(ptype, (24, 20, 16), d, a, 0.05, 0.5, seed=3)
    assert abs(float(m.astype(np.float64).mean())) < 1e-6 * 0.05
    assert abs(np.sqrt(np.mean(m.astype(np.float64) ** 2)) / 0.05 - 1) < 1e-5
    r, _ = R.spectral_ratio(m, ptype, d, a, 0.05, 0.5, 1e-6)
    assert np.max(np.abs(r - 1)) < 1e-3
