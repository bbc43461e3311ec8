"""Channel averaging (Core::uvshiftAndAverageBaselineFreq, core.cpp:1586-1802,
and Mode::averageFrequency, mode.cpp:1399-1424) in the oracle, pinned by
linearity: the averaged job's products are the group means of the same data
correlated without averaging (FP64 flavour), for averaging shorter than an
xmac stride and for averaging across several strides; zoom-band
autocorrelations start at the averaged parent's channel choffset/avg
(mode.cpp:396)."""
import ctypes as C

import numpy as np
import pytest

from fxcorr import JobSpec, check
from fxsynth import make_subint
from fxtestlib import oracle_results, plan_info


def off(orc, fn, job, *a):
    o = C.c_int64()
    check(orc, getattr(orc, fn)(C.byref(job), *a, C.byref(o)), fn)
    return o.value


@pytest.mark.parametrize("avg,zoom", [(4, [(0, 37, 64)]), (256, [])])
def test_averaged_products_are_group_means(oracle, avg, zoom):
    kw = dict(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=6, zoom=zoom)
    spec1, specn = JobSpec(**kw), JobSpec(chans_to_avg=avg, **kw)
    job1, k1 = spec1.build()
    jobn, kn = specn.build()
    sd = make_subint(spec1, seed=21)
    r1 = oracle_results(oracle, job1, sd, fp64=True)
    rn = oracle_results(oracle, jobn, sd, fp64=True)
    info = plan_info(oracle, "orc_", jobn)
    assert rn.size == info.result_length
    nfreqs = job1.nfreqs
    checked = 0
    for f in range(nfreqs):
        nchan = job1.freqs[f].nchan
        for b in range(job1.nbaselines):
            o1 = off(oracle, "orc_plan_baseline_offset", job1, f, b)
            on = off(oracle, "orc_plan_baseline_offset", jobn, f, b)
            if o1 < 0:
                continue
            npp = job1.baselines[b].npolproducts[0]
            for p in range(npp):
                full = r1[o1 + p * nchan: o1 + (p + 1) * nchan]
                want = full.reshape(-1, avg).mean(axis=1)
                got = rn[on + p * (nchan // avg): on + (p + 1) * (nchan // avg)]
                assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
                checked += 1
    assert checked > 0
    # autocorrelations: recorded bands then zoom bands, then the cross-pol pass
    for d in range(job1.ndatastreams):
        a1 = off(oracle, "orc_plan_autocorr_offset", job1, d)
        an = off(oracle, "orc_plan_autocorr_offset", jobn, d)
        ds = job1.datastreams[d]
        nrec, nz = ds.nrecbands, ds.nzoombands
        C_ = spec1.nchan
        for pas in range(2):
            base1 = a1 + pas * (nrec * C_ + sum(job1.freqs[ds.zoomband_freqindex[z]].nchan for z in range(nz)))
            i_n = an + pas * (nrec * C_ // avg + sum(job1.freqs[ds.zoomband_freqindex[z]].nchan // avg
                                                    for z in range(nz)))
            for k in range(nrec):
                full = r1[base1 + k * C_: base1 + (k + 1) * C_]
                got = rn[i_n: i_n + C_ // avg]
                assert np.allclose(got, full.reshape(-1, avg).mean(axis=1), rtol=1e-12)
                i_n += C_ // avg
            for z in range(nz):
                zn = job1.freqs[ds.zoomband_freqindex[z]].nchan // avg
                parent = ds.zoomband_parent[z]
                pfull = r1[base1 + parent * C_: base1 + (parent + 1) * C_].reshape(-1, avg).mean(axis=1)
                start = ds.zoomband_choffset[z] // avg
                got = rn[i_n: i_n + zn]
                assert np.allclose(got, pfull[start: start + zn], rtol=1e-12)
                i_n += zn
