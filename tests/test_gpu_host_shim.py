"""The C++ Mode/Core layer (difx-svn2git-try3_amd/host/gpumode.h) on the GPU.

tests/cpp/test_host_shim.cpp drives it the way mpifxcorr's Core drives Mode
(core.cpp:667-1045, 1299-1369): once through GpuCore::processdata for the
whole slot and once block by block through GpuMode::process with the
baseline and autocorrelation accumulation done on the host from getFreqs /
getConjugatedFreqs / getDataWeight / getAutocorrelation / getWeight.  Both
result buffers must match each other and the oracle to 1e-5 relative.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import fxcorr
from fxcorr import Baseline, Datastream, Freq, JobSpec
from fxsynth import make_subint
from fxtestlib import oracle_kurtosis, oracle_results, rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "difx-svn2git-try3_amd", "_build", "test_host_shim")


def write_case(d, job, sd, prec):
    hdr = np.zeros(16, dtype=np.int32)
    sub = sd.subint_desc()
    hdr[:10] = [job.nfreqs, job.ndatastreams, job.nbaselines, job.blockspersend, prec,
                job.crosspol_autocorrs, sd.scan_start_sec, sub.scan, sub.sec, sub.ns]
    with open(os.path.join(d, "job.bin"), "wb") as f:
        f.write(hdr.tobytes())
        f.write(C.string_at(job.freqs, C.sizeof(Freq) * job.nfreqs))
        f.write(C.string_at(job.datastreams, C.sizeof(Datastream) * job.ndatastreams))
        f.write(C.string_at(job.baselines, C.sizeof(Baseline) * job.nbaselines))
    for i, st in enumerate(sd.stations):
        st.raw.tofile(os.path.join(d, f"st{i}.raw"))
        vf = np.ascontiguousarray(st.validflags, dtype=np.int32)
        meta = np.concatenate([np.array([0, st.datasec, st.datans, vf.size], dtype=np.int32), vf,
                               np.zeros(vf.size & 1, dtype=np.int32)])
        with open(os.path.join(d, f"st{i}.meta"), "wb") as f:
            f.write(meta.tobytes())
            f.write(np.asarray(st.interpolator, dtype=np.float64).tobytes())


CASES = {
    "vdif": dict(spec=JobSpec(nstations=4, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=10),
                 invalid_frames={1: [3]}),
    "mark5b": dict(spec=JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=128, blockspersend=9,
                                fmt=fxcorr.FXG_FORMAT_MARK5B, framebytes=10016, headerbytes=16),
                   invalid_frames={}),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_core_and_mode_paths_agree_with_oracle(oracle, tmp_path, case):
    assert os.path.exists(EXE), "build the host shim test program first (__graft_entry__.build())"
    c = CASES[case]
    spec = c["spec"]
    job, keep = spec.build()
    sd = make_subint(spec, seed=11, invalid_frames=c["invalid_frames"])
    write_case(str(tmp_path), job, sd, fxcorr.FXG_XMAC_FP32)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = oracle_results(oracle, job, sd)
    core = np.fromfile(os.path.join(tmp_path, "results_core.bin"), dtype=np.complex64)
    blocks = np.fromfile(os.path.join(tmp_path, "results_blocks.bin"), dtype=np.complex64)
    assert core.size == ref.size == blocks.size
    assert rel_err(core, ref) <= 1e-5
    assert rel_err(blocks, ref) <= 1e-5
    # kurtosis records of two Core threads, 32 channels (GpuCore::setDumpKurtosis)
    k = np.fromfile(os.path.join(tmp_path, "kurtosis.bin"), dtype=np.float32)
    nb, nout = spec.nrecbands, min(32, spec.nchan)
    recs = k[:2 * spec.nstations * nb * nout].reshape(2, spec.nstations, nb, nout)
    valid = k[2 * spec.nstations * nb * nout:].reshape(2, spec.nstations)
    for d in range(spec.nstations):
        kref, kvalid = oracle_kurtosis(oracle, job, sd, d, 2, 32, nb, spec.nchan)
        np.testing.assert_array_equal(valid[:, d], kvalid)
        np.testing.assert_allclose(recs[:, d], kref, rtol=0, atol=5e-4)
