"""Result layout (Configuration::populateResultLengths, configuration.cpp:
2212-2441) on the reference's real job ma008_1 and on the benchmark shapes,
computed independently by the HIP library's planner and by the oracle; and
the C ABI of libfxcorr_gpu.so (loads and exports everything the header
declares, plans without a GPU, refuses to run without one)."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import fxcorr
from fxcorr import (Baseline, Datastream, Freq, JobDesc, JobSpec, PlanInfo, check, FXG_FORMAT_VDIF,
                    FXG_MAX_POLPRODUCTS)

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def ma008_job():
    t = json.load(open(os.path.join(HERE, "golden", "ma008_1_tables.json")))
    nf = len(t["freqs"])
    freqs = (Freq * nf)()
    for i, f in enumerate(t["freqs"]):
        freqs[i].freq_mhz, freqs[i].bw_mhz = f["freq_mhz"], f["bw_mhz"]
        freqs[i].lsb, freqs[i].nchan, freqs[i].chans_to_avg = int(f["lsb"]), f["nchan"], f["chans_to_avg"]

    def lowest(fi):
        f = t["freqs"][fi]
        return f["freq_mhz"] - (f["bw_mhz"] if f["lsb"] else 0.0)

    dss = (Datastream * len(t["datastreams"]))()
    for d, ds in enumerate(t["datastreams"]):
        x = dss[d]
        x.format, x.nbits, x.framebytes, x.headerbytes = FXG_FORMAT_VDIF, ds["nbits"], ds["framebytes"], 32
        x.nrecfreqs, x.nrecbands, x.nzoombands = len(ds["recfreqs"]), len(ds["recbands"]), len(ds["zoombands"])
        for i, rf in enumerate(ds["recfreqs"]):
            x.recfreq_index[i] = rf["index"]
            x.recfreq_clockoffset_us[i] = rf["clock_us"]
            x.recfreq_looffset_hz[i] = rf["looffset_hz"]
        pols = bytearray(fxcorr.FXG_MAX_BANDS)
        for b, rb in enumerate(ds["recbands"]):
            x.recband_freq[b] = rb["localfreq"]
            pols[b] = ord(rb["pol"])
        C.memmove(C.addressof(x) + Datastream.recband_pol.offset, bytes(pols), len(pols))
        zp = bytearray(fxcorr.FXG_MAX_ZOOMBANDS)
        for z, zb in enumerate(ds["zoombands"]):
            fz = ds["zoomfreqs"][zb["localfreq"]]["index"]
            x.zoomband_freqindex[z] = fz
            zp[z] = ord(zb["pol"])
            # parent: the recorded band of the same polarisation containing the zoom band
            for b, rb in enumerate(ds["recbands"]):
                fp = ds["recfreqs"][rb["localfreq"]]["index"]
                chbw = t["freqs"][fp]["bw_mhz"] / t["freqs"][fp]["nchan"]
                off = (lowest(fz) - lowest(fp)) / chbw
                if rb["pol"] == zb["pol"] and off >= -1e-6 and \
                        off + t["freqs"][fz]["nchan"] <= t["freqs"][fp]["nchan"] + 1e-6:
                    x.zoomband_parent[z] = b
                    x.zoomband_choffset[z] = int(round(off))
                    break
            else:
                raise AssertionError("zoom band without parent")
        C.memmove(C.addressof(x) + Datastream.zoomband_pol.offset, bytes(zp), len(zp))
    bls = (Baseline * len(t["baselines"]))()
    for b, bl in enumerate(t["baselines"]):
        x = bls[b]
        x.ds1, x.ds2, x.nfreqs = bl["ds1"], bl["ds2"], len(bl["freqs"])
        ds1 = t["datastreams"][bl["ds1"]]
        for f, prods in enumerate(bl["freqs"]):
            b1 = prods[0][0]
            nrec = len(ds1["recbands"])
            if b1 < nrec:
                fi = ds1["recfreqs"][ds1["recbands"][b1]["localfreq"]]["index"]
            else:
                fi = ds1["zoomfreqs"][ds1["zoombands"][b1 - nrec]["localfreq"]]["index"]
            x.freq_index[f] = fi
            x.npolproducts[f] = len(prods)
            for p, (a, c) in enumerate(prods):
                x.band1[f][p], x.band2[f][p] = a, c
    job = JobDesc()
    job.nfreqs, job.ndatastreams, job.nbaselines = nf, len(t["datastreams"]), len(t["baselines"])
    job.blockspersend, job.fringe_order = 16, 1
    job.crosspol_autocorrs = 1
    job.freqs = C.cast(freqs, C.POINTER(Freq))
    job.datastreams = C.cast(dss, C.POINTER(Datastream))
    job.baselines = C.cast(bls, C.POINTER(Baseline))
    return job, (freqs, dss, bls), t


def offsets(lib, prefix, job, sample=None):
    info = PlanInfo()
    check(lib, getattr(lib, prefix + "plan")(C.byref(job), C.byref(info)), "plan")
    o = C.c_int64()
    bl, bw = {}, {}
    pairs = [(f, b) for f in range(job.nfreqs) for b in range(job.nbaselines)]
    if sample is not None and len(pairs) > sample:
        rng = np.random.default_rng(0)
        pairs = [pairs[i] for i in sorted(rng.choice(len(pairs), sample, replace=False))] + [pairs[0], pairs[-1]]
    for f, b in pairs:
        if True:
            check(lib, getattr(lib, prefix + "plan_baseline_offset")(C.byref(job), f, b, C.byref(o)), "bl")
            bl[f, b] = o.value
            check(lib, getattr(lib, prefix + "plan_bweight_offset")(C.byref(job), f, b, C.byref(o)), "bw")
            bw[f, b] = o.value
    ac, acw = [], []
    for d in range(job.ndatastreams):
        check(lib, getattr(lib, prefix + "plan_autocorr_offset")(C.byref(job), d, C.byref(o)), "ac")
        ac.append(o.value)
        check(lib, getattr(lib, prefix + "plan_acweight_offset")(C.byref(job), d, C.byref(o)), "acw")
        acw.append(o.value)
    return info, bl, bw, ac, acw


def region_sizes(job, info):
    """The size of every region from its meaning, in layout order."""
    used = set()
    for b in range(job.nbaselines):
        for f in range(job.baselines[b].nfreqs):
            used.add(job.baselines[b].freq_index[f])
    x = bw = 0
    for f in sorted(used):
        for b in range(job.nbaselines):
            bl = job.baselines[b]
            for lf in range(bl.nfreqs):
                if bl.freq_index[lf] == f:
                    x += bl.npolproducts[lf] * job.freqs[f].nchan
                    bw += (bl.npolproducts[lf] + 1) // 2
    return x, bw, used


def test_ma008_layout_two_implementations_agree(oracle):
    lib = fxcorr.load_gpu_lib()
    job, keep, t = ma008_job()
    a = offsets(lib, "fxg_", job)
    b = offsets(oracle, "orc_", job)
    assert a[0].result_length == b[0].result_length
    assert a[1:] == b[1:]
    info, bl, bw, ac, acw = a
    x, nbw, used = region_sizes(job, info)
    # cross-correlations first, from offset 0, frequency-major then baseline
    assert min(v for v in bl.values() if v >= 0) == 0
    assert min(v for v in bw.values() if v >= 0) == x
    assert ac[0] == x + nbw
    # autocorrelation regions are contiguous and ordered by datastream
    assert all(ac[d] < ac[d + 1] for d in range(len(ac) - 1))
    assert acw[0] > ac[-1]
    # every used frequency gets its autocorrelations: 2 passes (maxproducts 4)
    assert info.bandsperautocorr == 2
    total_ac = 0
    for d in range(job.ndatastreams):
        ds = job.datastreams[d]
        for k in range(ds.nrecbands + ds.nzoombands):
            fi = ds.recfreq_index[ds.recband_freq[k]] if k < ds.nrecbands else ds.zoomband_freqindex[k - ds.nrecbands]
            if fi in used:
                total_ac += 2 * job.freqs[fi].nchan
    assert acw[0] - ac[0] >= total_ac
    assert info.xmacstridelen == 128          # gcd of the correlated channel counts is 128 -> 128 (configuration.cpp:2491)


@pytest.mark.parametrize("shape", ["c1", "c2", "c3"])
def test_config_layouts_agree(oracle, shape):
    lib = fxcorr.load_gpu_lib()
    spec = {"c1": JobSpec(nstations=4, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256),
            "c2": JobSpec(nstations=10, nsubbands=8, npol=2, bw_mhz=32.0, nchan=1024),
            "c3": JobSpec(nstations=32, nsubbands=16, npol=2, bw_mhz=64.0, nchan=4096)}[shape]
    job, keep = spec.build()
    a = offsets(lib, "fxg_", job, sample=12)
    b = offsets(oracle, "orc_", job, sample=12)
    assert a[0].result_length == b[0].result_length and a[1:] == b[1:]
    nbl = spec.nstations * (spec.nstations - 1) // 2
    x = nbl * spec.nsubbands * 4 * spec.nchan
    w = nbl * spec.nsubbands * 2
    acs = spec.nstations * spec.nrecbands * 2 * spec.nchan
    acw = spec.nstations * ((spec.nrecbands * 2 + 1) // 2)
    assert a[0].result_length == x + w + acs + acw


def declared_symbols():
    hdr = open(os.path.join(REPO, "include", "fxcorr_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fxg_\w+)\s*\(", hdr, re.M)))


def test_library_exports_every_declared_symbol():
    lib = fxcorr.load_gpu_lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.fxg_abi_version() == 3


def test_host_library_exports_serving_abi():
    host = fxcorr.load_host_lib()
    hdr = open(os.path.join(REPO, "include", "fxcorr_serve.h")).read()
    syms = sorted(set(re.findall(r"^\s*int\s+(fxs_\w+)\s*\(", hdr, re.M)))
    assert len(syms) == 7
    for s in syms:
        assert hasattr(host, s), s


def test_plan_without_gpu_and_error_paths():
    lib = fxcorr.load_gpu_lib()
    job, keep = JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256).build()
    info = PlanInfo()
    assert lib.fxg_plan(C.byref(job), C.byref(info)) == 0
    assert info.result_length > 0 and info.device_bytes > 0
    bad, keep2 = JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256).build()
    bad.fringe_order = 3
    assert lib.fxg_plan(C.byref(bad), C.byref(info)) == -4
    assert b"order" in lib.fxg_last_error()
    bad.fringe_order = 1
    bad.baselines[0].band1[0][0] = 99
    assert lib.fxg_plan(C.byref(bad), C.byref(info)) == -1


def test_complex_sampling_plan_rules():
    """Complex sampling: fftchannels = nchan; what the reference refuses
    (post-F fringe rotation) or this build does not take (mixed real and
    complex datastreams) is refused at planning, by the library and by the
    oracle alike.  One band of 2-bit complex samples and phase cal of complex
    data plan, with the same result layout in both."""
    lib, orc = fxcorr.load_gpu_lib(), fxcorr.load_oracle()
    info = PlanInfo()
    job, keep = JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, sampling="complex").build()
    assert lib.fxg_plan(C.byref(job), C.byref(info)) == 0 and info.fftchannels == 256
    ref = PlanInfo()
    assert orc.orc_plan(C.byref(job), C.byref(ref)) == 0 and ref.result_length == info.result_length
    job.fringe_order = 0
    assert lib.fxg_plan(C.byref(job), C.byref(info)) == -4
    job.fringe_order = 1
    job.datastreams[1].sampling = 0
    assert lib.fxg_plan(C.byref(job), C.byref(info)) == -4
    one, k3 = JobSpec(nstations=2, nsubbands=1, npol=1, bw_mhz=16.0, nchan=256, sampling="complex").build()
    assert lib.fxg_plan(C.byref(one), C.byref(info)) == 0
    assert orc.orc_plan(C.byref(one), C.byref(ref)) == 0 and ref.result_length == info.result_length
    pc, k4 = JobSpec(nstations=2, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, sampling="complex",
                     pcal_interval_mhz=1.0).build()
    assert lib.fxg_plan(C.byref(pc), C.byref(info)) == 0
    assert orc.orc_plan(C.byref(pc), C.byref(ref)) == 0 and ref.result_length == info.result_length
    nt, off = C.c_int32(), C.c_int32()
    assert lib.fxg_plan_pcal_tones(C.byref(pc), 0, 0, C.byref(nt), C.byref(off)) == 0
    assert (nt.value, off.value) == (16, 1000000)


def test_context_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = fxcorr.load_gpu_lib()
    ctx = C.c_void_p()
    assert lib.fxg_ctx_create(0, C.byref(ctx)) == -6
