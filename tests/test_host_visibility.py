"""Visibility (mpifxcorr/src/visibility.cpp:181-193, 348-560, 716-900,
1050-1131) as restated by difx-svn2git-try3_amd/host/visibility.cpp, on
oracle results; no GPU needed.

The output file is read back the way difxio's DifxVisRecordgetnext reads
binary records (libraries/difxio/difxio/parsevis.c:126-197): sync word,
header version, then baseline, mjd, seconds, config, source, freq, polpair,
pulsar bin, weight, uvw, and nchan complex floats.  The normalisation is
pinned by what it means:
  - correlation coefficients: identical voltages at two stations give a
    cross product equal to either autocorrelation, and the autocorrelation
    of white noise is the quantiser's mean power (2-bit levels +-1, +-3.3359
    with thresholds at +-0.98 sigma: 4.313);
  - Tsys mode: each autocorrelation's band mean becomes Tsys, and the
    identical-voltage cross product's band mean becomes Tsys over the 2-bit
    decorrelation factor 0.88 (mode.cpp:1594).
"""
import ctypes as C
import math
import os
import struct
import subprocess

import numpy as np
import pytest

from fxcorr import Baseline, Datastream, Freq, JobSpec
from fxsynth import make_subint
from fxtestlib import oracle_results

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "difx-svn2git-try3_amd", "_build", "test_visibility")
SYNC = 0xFF00FF00


def read_difx(path, job):
    """Records as difxio's parsevis reads them; nchan from the freq table."""
    recs = []
    with open(path, "rb") as f:
        buf = f.read()
    i = 0
    while i < len(buf):
        sync, ver = struct.unpack_from("<Ii", buf, i)
        assert sync == SYNC and ver == 1, (hex(sync), ver)
        i += 8
        bl, mjd = struct.unpack_from("<ii", buf, i); i += 8
        sec, = struct.unpack_from("<d", buf, i); i += 8
        cfg, src, fq = struct.unpack_from("<iii", buf, i); i += 12
        pp = buf[i:i + 2].decode(); i += 2
        pbin, = struct.unpack_from("<i", buf, i); i += 4
        w, = struct.unpack_from("<d", buf, i); i += 8
        uvw = struct.unpack_from("<ddd", buf, i); i += 24
        nch = job.freqs[fq].nchan // job.freqs[fq].chans_to_avg
        data = np.frombuffer(buf, dtype=np.complex64, count=nch, offset=i); i += 8 * nch
        recs.append(dict(baseline=bl, mjd=mjd, seconds=sec, config=cfg, source=src, freq=fq, pol=pp,
                         bin=pbin, weight=w, uvw=uvw, data=data))
    return recs


def write_job(d, job, sd):
    hdr = np.zeros(16, dtype=np.int32)
    hdr[:7] = [job.nfreqs, job.ndatastreams, job.nbaselines, job.blockspersend, 0, job.crosspol_autocorrs, 0]
    with open(os.path.join(d, "job.bin"), "wb") as f:
        f.write(hdr.tobytes())
        f.write(C.string_at(job.freqs, C.sizeof(Freq) * job.nfreqs))
        f.write(C.string_at(job.datastreams, C.sizeof(Datastream) * job.ndatastreams))
        f.write(C.string_at(job.baselines, C.sizeof(Baseline) * job.nbaselines))


def run_case(oracle, tmp_path, tsys, nsub=3):
    spec = JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=128, blockspersend=40)
    job, keep = spec.build()
    sd = make_subint(spec, seed=31, rho=0.0, delays_us=[0.0, 0.0, 2.0])
    sd.stations[1].raw = sd.stations[0].raw.copy()            # identical voltages, identical models
    sd.stations[1].interpolator = sd.stations[0].interpolator.copy()
    res = oracle_results(oracle, job, sd)
    write_job(str(tmp_path), job, sd)
    res.tofile(os.path.join(tmp_path, "results.bin"))
    subintns = 8_000_000
    r = subprocess.run([EXE, str(tmp_path), str(nsub), str(tsys), str(subintns), str(nsub * subintns / 1e9)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"done=1 subintsthisintegration={nsub}" in r.stdout
    return spec, job, read_difx(os.path.join(tmp_path, "out.difx"), job)


def test_records_and_coefficients(oracle, tmp_path):
    assert os.path.exists(EXE), "build first (__graft_entry__.build())"
    spec, job, recs = run_case(oracle, tmp_path, 0.0)
    nbl = 3
    ncross = nbl * spec.nsubbands * 4
    nauto = 3 * spec.nrecbands * 2                            # parallel and cross-pol passes
    assert len(recs) == ncross + nauto
    tel = [1, 3, 5]
    for r in recs:
        assert r["mjd"] == 57000 + (86000 + 500 + 7) // 86400
        assert r["seconds"] == (86000 + 500 + 7) % 86400 + 0.25 + 0.012
        assert r["config"] == 0 and r["source"] == 3 and r["bin"] == 0
        assert abs(r["weight"] - 1.0) < 1e-6                  # every block valid
    cross = [r for r in recs if r["baseline"] % 257 != 0 or r["baseline"] // 257 == 0]
    autos = [r for r in recs if r not in cross]
    assert len(cross) == ncross
    assert {r["baseline"] for r in cross} == {(tel[a] + 1) * 256 + tel[b] + 1 for a, b in ((0, 1), (0, 2), (1, 2))}
    assert {r["baseline"] for r in autos} == {257 * (t + 1) for t in tel}
    # white-noise autocorrelation = mean quantiser power
    p_in = math.erf(0.98 / math.sqrt(2))
    power = p_in * 1.0 + (1 - p_in) * 3.3359 ** 2
    for r in autos:
        if r["pol"][0] == r["pol"][1]:
            v = r["data"]
            assert abs(v.real[1:].mean() / power - 1) < 0.02, v.real.mean()
            assert np.max(np.abs(v.imag)) < 1e-3 * power
    # identical voltages: the cross product equals the autocorrelation
    a0 = {(r["freq"], r["pol"]): r["data"] for r in autos if r["baseline"] == 257 * 2}
    for r in cross:
        if r["baseline"] == 2 * 256 + 4 and r["pol"][0] == r["pol"][1]:
            np.testing.assert_allclose(r["data"], a0[(r["freq"], r["pol"])], rtol=2e-5, atol=1e-5)


def test_tsys_calibration(oracle, tmp_path):
    tsys = 57.5
    spec, job, recs = run_case(oracle, tmp_path, tsys)
    for r in recs:
        v = r["data"]
        if r["baseline"] % 257 == 0 and r["pol"][0] == r["pol"][1]:
            assert abs(v.real.mean() / tsys - 1) < 1e-5
        if r["baseline"] == 2 * 256 + 4 and r["pol"][0] == r["pol"][1]:
            assert abs(v.real.mean() / (tsys / 0.88) - 1) < 1e-4
