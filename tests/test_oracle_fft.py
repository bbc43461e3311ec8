"""Pins the oracle's FFT (the reference calls FFTW's unnormalised forward
transform, architecture.h.in:825-838, 1024-1028) against numpy's pocketfft,
an independent library routine, and against closed forms."""
import ctypes as C

import numpy as np
import pytest


def orc_fft(orc, x, fp64):
    dt = np.complex128 if fp64 else np.complex64
    a = np.ascontiguousarray(x, dtype=dt)
    out = np.zeros_like(a)
    assert orc.orc_fft(a.size, 1 if fp64 else 0, a.ctypes.data, out.ctypes.data) == 0
    return out


@pytest.mark.parametrize("L", [16, 128, 512, 2048, 8192])
def test_fft_vs_numpy(oracle, L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    ref = np.fft.fft(x)
    e64 = np.linalg.norm(orc_fft(oracle, x, True) - ref) / np.linalg.norm(ref)
    e32 = np.linalg.norm(orc_fft(oracle, x, False) - ref) / np.linalg.norm(ref)
    assert e64 < 1e-13
    assert e32 < 2e-7 * np.log2(L)


def test_fft_closed_forms(oracle):
    L = 1024
    imp = np.zeros(L, complex); imp[0] = 1.0
    assert np.allclose(orc_fft(oracle, imp, True), np.ones(L), atol=1e-15)
    k0 = 37
    tone = np.exp(2j * np.pi * k0 * np.arange(L) / L)
    X = orc_fft(oracle, tone, True)
    assert abs(X[k0] - L) < 1e-9 and np.max(np.abs(np.delete(X, k0))) < 1e-9
    # Parseval
    rng = np.random.default_rng(0)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    X = orc_fft(oracle, x, True)
    assert abs(np.sum(np.abs(X) ** 2) - L * np.sum(np.abs(x) ** 2)) < 1e-8 * L * np.sum(np.abs(x) ** 2)
