"""Spectral kurtosis in the oracle (Mode's dumpkurtosis path,
mpifxcorr/src/mode.cpp:1260-1277 and 1426-1478, with Core's thread block
ranges core.cpp:2089-2095), pinned without the reference binary:

* by an independent FP64 restatement here, from the oracle's FP64 spectra,
  of the spectral-kurtosis estimator (n/(n-1)) (n sum|X|^4 / (sum|X|^2)^2 - 1)
  - 1 over each thread's blocks and of the channel sums; a wrong n, a
  missing correction factor, a wrong block range or an averaging mistake
  moves it by far more than FP32 rounding;
* by what the mathematics fixes: a constant-modulus (CW tone) channel has
  sk = -1; Gaussian noise has sk = 0 on average (the estimator is unbiased);
* by the validity rule: a station without data has no nonzero sums.
"""
import numpy as np
import pytest

from fxcorr import JobSpec
from fxsynth import make_subint
from fxtestlib import oracle_kurtosis, oracle_spectra


def small_spec(**kw):
    base = dict(nstations=2, nsubbands=2, npol=2, bw_mhz=16.0, nchan=128, blockspersend=10, crosspol=False)
    base.update(kw)
    return JobSpec(**base)


def fp64_kurtosis(orc, job, sd, spec, ds, nthreads, maxch):
    """The estimator in FP64 from the oracle's FP64 spectra (the fractional
    rotation has unit modulus, so |X| is that of the unrotated spectrum)."""
    n, nb, C = spec.blockspersend, spec.nrecbands, spec.nchan
    spectra = np.stack([oracle_spectra(orc, job, sd, ds, b, nb, C, fp64=True) for b in range(n)])
    p = np.abs(spectra) ** 2                                  # [block][band][chan]
    out, valid = [], []
    start = 0
    for t in range(nthreads):
        cnt = n // nthreads + (1 if t < n % nthreads else 0)
        s1 = p[start:start + cnt].sum(axis=0)
        s2 = (p[start:start + cnt] ** 2).sum(axis=0)
        start += cnt
        with np.errstate(divide="ignore", invalid="ignore"):
            sk = (cnt / (cnt - 1.0)) * (cnt * s2 / s1 ** 2 - 1.0) - 1.0
        nz = (s1 > 0).any(axis=1)
        sk[~nz] = 0.0
        if maxch < C:
            k = C // maxch
            sk = sk[:, :maxch * k].reshape(nb, maxch, k).mean(axis=2)
        out.append(sk)
        valid.append(int(nz[-1]))
    return np.stack(out), np.array(valid)


@pytest.mark.parametrize("nthreads,maxch", [(1, 128), (3, 128), (3, 32), (4, 1)])
def test_kurtosis_matches_fp64_estimator(oracle, nthreads, maxch):
    spec = small_spec()
    job, keep = spec.build()
    sd = make_subint(spec, seed=11, rho=0.05)
    for ds in range(spec.nstations):
        got, valid = oracle_kurtosis(oracle, job, sd, ds, nthreads, maxch, spec.nrecbands, spec.nchan)
        ref, rvalid = fp64_kurtosis(oracle, job, sd, spec, ds, nthreads, maxch)
        assert got.shape == ref.shape
        np.testing.assert_array_equal(valid, rvalid)
        assert np.all(valid == 1)
        # FP32 sums of 2..10 blocks against FP64: ~1e-6 relative in the sums
        np.testing.assert_allclose(got, ref, rtol=0, atol=2e-4)


def test_kurtosis_thread_ranges_differ(oracle):
    """Records of different threads cover different blocks: swapping the
    range rule (e.g. n*t/T) changes them."""
    spec = small_spec(blockspersend=10)
    job, keep = spec.build()
    sd = make_subint(spec, seed=3)
    got, _ = oracle_kurtosis(oracle, job, sd, 0, 3, 128, spec.nrecbands, spec.nchan)
    assert not np.allclose(got[0], got[1]) and not np.allclose(got[1], got[2])


def test_kurtosis_cw_tone_is_minus_one(oracle):
    """A strong CW tone has constant modulus in its channel: sk = -1 there,
    about 0 elsewhere (Gaussian noise)."""
    spec = small_spec(nstations=1, nsubbands=1, npol=1, nbits=8, blockspersend=40)
    job, keep = spec.build()
    fs = 2e6 * spec.bw_mhz
    chan = 37
    f = chan * fs / spec.fftchannels                  # on channel 37: no leakage into the others
    sd = make_subint(spec, seed=5, rho=0.0, delays_us=[0.0], rates_us_per_s=[0.0],
                     pcal=dict(amp=10.0, spacing_hz=fs, offsets_hz=[f], ntones=[1]))
    got, valid = oracle_kurtosis(oracle, job, sd, 0, 1, 1 << 20, spec.nrecbands, spec.nchan)
    sk = got[0, 0]
    assert valid[0] == 1
    assert abs(sk[chan] + 1.0) < 2e-3, sk[chan]
    others = np.delete(sk, [0, chan - 1, chan, chan + 1])
    # 40 blocks: each channel's sk has a standard deviation of about 0.3
    assert abs(others.mean()) < 0.1, others.mean()
    assert others.std() < 0.6


def test_kurtosis_gaussian_noise_unbiased(oracle):
    spec = small_spec(nstations=1, nsubbands=2, npol=2, blockspersend=64)
    job, keep = spec.build()
    sd = make_subint(spec, seed=21, rho=0.0)
    got, valid = oracle_kurtosis(oracle, job, sd, 0, 1, 1 << 20, spec.nrecbands, spec.nchan)
    sk = got[0][:, 1:]                                 # the DC channel carries the 2-bit bias
    assert abs(sk.mean()) < 0.05, sk.mean()
    # Var(sk) ~ 4/n for complex Gaussian channels
    assert 0.5 * 2 / np.sqrt(64) < sk.std() < 2.0 * 2 / np.sqrt(64), sk.std()


def test_kurtosis_invalid_station(oracle):
    """A station whose data cannot be used: Mode::process zeroes every block
    and returns before the accumulation, so no record is valid."""
    spec = small_spec()
    job, keep = spec.build()
    sd = make_subint(spec, seed=2, invalid_blocks={1: list(range(spec.blockspersend))})
    got, valid = oracle_kurtosis(oracle, job, sd, 1, 2, 128, spec.nrecbands, spec.nchan)
    assert np.all(valid == 0)
    assert np.all(got == 0)
    got0, valid0 = oracle_kurtosis(oracle, job, sd, 0, 2, 128, spec.nrecbands, spec.nchan)
    assert np.all(valid0 == 1)
