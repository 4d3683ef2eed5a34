"""Pins the oracle's 3-D real FFT (oracle/kapsel_oracle.cpp fft_r2c/fft_c2r).

Checks: the reference's own Ooura rdft3d + rdft3dsort (golden vectors in
tests/golden, and oracle/_ref when built), numpy.fft.rfftn/irfftn, and the
DFT's closed forms (delta, single cosine, Parseval, round trip).
"""
import glob
import os

import numpy as np
import pytest

from icep_kapsel_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))


def _oracle(shape):
    from oracle.oracle import Oracle
    return Oracle(abi.make_params(shape))


def _pad(real, nz):
    nx, ny = real.shape[:2]
    a = np.zeros((nx, ny, nz + 2))
    a[:, :, :nz] = real
    return a


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "ooura_fft_*.npz"))))
def test_oracle_fft_matches_reference_golden(path):
    g = np.load(path)
    nx, ny, nz = g["real_in"].shape
    o = _oracle((nx, ny, nz))
    o.set_field("U0", _pad(g["real_in"], nz))
    o.fft_r2c("U0")
    spec = o.get_field("U0")
    ref = g["spec_out"]
    assert np.abs(spec - ref).max() <= 1e-13 * np.abs(ref).max()
    o.set_field("U1", g["spec_in"])
    o.fft_c2r("U1")
    back = o.get_field("U1")[:, :, :nz]
    assert np.abs(back - g["real_out"]).max() <= 1e-14 * max(1.0, np.abs(g["real_out"]).max())


def test_oracle_fft_matches_reference_build():
    from oracle.oracle import REF_LIB, OouraRef
    if not os.path.exists(REF_LIB):
        pytest.skip("reference Ooura FFT not built (needs /root/reference)")
    ref = OouraRef()
    rng = np.random.default_rng(5)
    for shape in [(16, 32, 64), (64, 64, 64)]:
        nx, ny, nz = shape
        a = _pad(rng.standard_normal(shape), nz)
        o = _oracle(shape)
        o.set_field("U0", a)
        o.fft_r2c("U0")
        r = ref.a2a_k(a)
        assert np.abs(o.get_field("U0") - r).max() <= 1e-13 * np.abs(r).max()


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 64), (64, 32, 16)])
def test_oracle_fft_matches_numpy(shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(11)
    real = rng.standard_normal(shape)
    o = _oracle(shape)
    o.set_field("U0", _pad(real, nz))
    o.fft_r2c("U0")
    spec = o.get_field("U0").view(np.complex128)
    ref = np.fft.rfftn(real)
    assert np.abs(spec - ref).max() <= 1e-13 * np.abs(ref).max()
    o.fft_c2r("U0")
    back = o.get_field("U0")[:, :, :nz]
    assert np.abs(back - real).max() <= 1e-13


def test_oracle_fft_closed_forms():
    n = 32
    o = _oracle(n)
    # delta at the origin: constant unit spectrum
    a = np.zeros((n, n, n))
    a[0, 0, 0] = 1.0
    o.set_field("U0", _pad(a, n))
    o.fft_r2c("U0")
    spec = o.get_field("U0").view(np.complex128)
    assert np.allclose(spec, 1.0, atol=1e-15)
    # delta at x0: e^{-i k.x0}, forward sign of A2a_k
    x0 = (3, 5, 7)
    a = np.zeros((n, n, n))
    a[x0] = 1.0
    o.set_field("U0", _pad(a, n))
    o.fft_r2c("U0")
    spec = o.get_field("U0").view(np.complex128)
    k = np.fft.fftfreq(n, 1.0 / n)
    kz = np.arange(n // 2 + 1)
    ph = np.exp(-2j * np.pi * (k[:, None, None] * x0[0] + k[None, :, None] * x0[1] + kz[None, None, :] * x0[2]) / n)
    assert np.abs(spec - ph).max() < 1e-13
    # single cosine: two conjugate spikes of height N^3/2 (one in the half spectrum + mirror)
    m = (2, 3, 4)
    x = np.arange(n)
    cosf = np.cos(2 * np.pi * (m[0] * x[:, None, None] + m[1] * x[None, :, None] + m[2] * x[None, None, :]) / n)
    o.set_field("U0", _pad(cosf, n))
    o.fft_r2c("U0")
    spec = o.get_field("U0").view(np.complex128)
    assert abs(spec[m] - n ** 3 / 2) < 1e-9
    spec[m] = 0
    assert np.abs(spec).max() < 1e-9
    # Parseval over the half spectrum
    rng = np.random.default_rng(3)
    real = rng.standard_normal((n, n, n))
    o.set_field("U0", _pad(real, n))
    o.fft_r2c("U0")
    spec = o.get_field("U0").view(np.complex128)
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    lhs = np.sum(real ** 2)
    rhs = np.sum(w * np.abs(spec) ** 2) / n ** 3
    assert abs(lhs - rhs) < 1e-10 * lhs
