"""GPU parity at the benchmark workloads' own shapes.

  - 32 stations x 2 pol x 16 subbands x 64 MHz, 2-bit VDIF, 4096 channels
    (8192-point transforms, 64-input cross-multiply on the matrix cores),
    4 blocks with a curved delay model (delay rates up to 3 us/s and a
    quadratic term), an invalid frame and a flagged block: FP32 within 1e-5
    of the FP32 oracle, fp16 within 1e-3 and bf16 within 8e-3 of the FP64
    oracle (||dV||/||V||; the spectra entering the matrix cores carry an
    11-bit or 8-bit significand);
  - the serving workload's three jobs (10, 20, 32 stations; 1024, 4096, 8192
    channels; 8 subbands x 64 MHz, 2-bit, fp16 cross-multiply), each against
    the FP64 oracle, and each equal bit for bit alone and in the mix of the
    request loop (include/fxcorr_serve.h);
  - the six uneven jobs (4..30 stations, 64..32768 channels, 2/8-bit, zoom
    bands) in one graph launch (fxg_batch_process): the batch equals each job
    alone bit for bit, and each job matches the oracle;
  - the DiFX writer (Visibility::addData / writedata, host/visibility.cpp)
    fed with the HIP library's results writes the records it writes from the
    FP32 oracle's results of the same subintegrations, within 1e-5.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import fxcorr
from fxcorr import JobSpec, check
from fxsynth import make_subint, uneven_workload_specs
from fxtestlib import GpuJob, oracle_results, plan_info, region_slices, rel_err

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5
MIXED_TOL = {"fp16": 1e-3, "bf16": 8e-3}
THREADS = max(1, min(16, os.cpu_count() or 1))


def threads_for(spec):
    """Oracle threads: each holds a whole result buffer, so no more than blocks."""
    return max(1, min(THREADS, spec.blockspersend))


def flush(msg):
    print(msg, flush=True)


# ---------------------------------------------------------------------------
# 32 stations x 16 subbands x 4096 channels
# ---------------------------------------------------------------------------
def c3_spec(xmac):
    return JobSpec(nstations=32, nsubbands=16, npol=2, bw_mhz=64.0, nchan=4096, nbits=2, blockspersend=4,
                   crosspol=True, xmac=xmac, framebytes=8032, base_freq_mhz=8200.0)


_C3 = {}


def c3_case():
    """The subintegration and its FP32 / FP64 oracle results (computed once)."""
    if not _C3:
        rng = np.random.default_rng(77)
        spec = c3_spec("fp32")
        sd = make_subint(spec, seed=77, rho=list(10.0 ** rng.uniform(-4.0, -1.0, spec.nsubbands)),
                         delays_us=rng.uniform(-20.0, 20.0, spec.nstations),
                         rates_us_per_s=rng.uniform(-3.0, 3.0, spec.nstations),
                         invalid_frames={3: [2]}, invalid_blocks={5: [1]})
        # a quadratic term in every station's model (us per block^2)
        for i, st in enumerate(sd.stations):
            st.interpolator = st.interpolator + np.array([2e-6 * (i - 15.5), 0.0, 0.0])
        job, keep = spec.build()
        flush("c3: FP32 oracle")
        _C3["ref32"] = oracle_results(oracle_lib(), job, sd, nthreads=threads_for(spec))
        flush("c3: FP64 oracle")
        _C3["ref64"] = oracle_results(oracle_lib(), job, sd, fp64=True, nthreads=threads_for(spec))
        _C3["sd"] = sd
        _C3["regions"] = region_slices(oracle_lib(), job, plan_info(oracle_lib(), "orc_", job))
    return _C3


def oracle_lib():
    return fxcorr.load_oracle()


@pytest.mark.parametrize("xmac", ["fp32", "fp16", "bf16"])
def test_32_station_workload_against_oracle(gpu, xmac):
    lib, ctx = gpu
    case = c3_case()
    spec = c3_spec(xmac)
    job, keep = spec.build()
    flush(f"c3: GPU {xmac}")
    g = GpuJob(lib, ctx, job)
    got = g.process(case["sd"])
    g.close()
    xend, a0, a1 = case["regions"]
    ref = case["ref32"] if xmac == "fp32" else case["ref64"]
    tol = FP32_TOL if xmac == "fp32" else MIXED_TOL[xmac]
    ex, ea = rel_err(got[:xend], ref[:xend]), rel_err(got[a0:a1], ref[a0:a1])
    flush(f"c3 {xmac}: cross {ex:.3e} auto {ea:.3e}")
    assert ex < tol and ea < tol, (ex, ea)
    # baseline and autocorrelation weights: sums of the same FP32 terms
    assert np.array_equal(got[xend:a0].view(np.float32), case["ref32"][xend:a0].view(np.float32))
    assert np.array_equal(got[a1:].view(np.float32), case["ref32"][a1:].view(np.float32))


# ---------------------------------------------------------------------------
# the serving workload's three jobs
# ---------------------------------------------------------------------------
SERVING = [(10, 1024), (20, 4096), (32, 8192)]


def serving_spec(nst, nchan, blocks=3):
    return JobSpec(nstations=nst, nsubbands=8, npol=2, bw_mhz=64.0, nchan=nchan, nbits=2, blockspersend=blocks,
                   xmac="fp16", base_freq_mhz=8200.0)


def serving_data(i, nst, nchan, k=0):
    spec = serving_spec(nst, nchan)
    rng = np.random.default_rng(300 + 10 * i + k)
    return spec, make_subint(spec, seed=300 + 10 * i + k, rho=0.02, delays_us=rng.uniform(-20.0, 20.0, nst),
                             rates_us_per_s=rng.uniform(-3.0, 3.0, nst), invalid_frames={0: [1 + k]})


@pytest.mark.parametrize("i", range(len(SERVING)), ids=[f"{n}st_{c}ch" for n, c in SERVING])
def test_serving_jobs_against_oracle(gpu, oracle, i):
    lib, ctx = gpu
    nst, nchan = SERVING[i]
    spec, sd = serving_data(i, nst, nchan)
    job, keep = spec.build()
    flush(f"serving job {nst}st {nchan}ch: GPU")
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    g.close()
    flush(f"serving job {nst}st {nchan}ch: FP64 oracle")
    ref = oracle_results(oracle, job, sd, fp64=True, nthreads=threads_for(spec))
    xend, a0, a1 = region_slices(oracle, job, plan_info(oracle, "orc_", job))
    ex, ea = rel_err(got[:xend], ref[:xend]), rel_err(got[a0:a1], ref[a0:a1])
    flush(f"  cross {ex:.3e} auto {ea:.3e}")
    assert ex < MIXED_TOL["fp16"] and ea < MIXED_TOL["fp16"], (ex, ea)


def test_serving_jobs_alone_equal_in_the_mix(gpu):
    """The three serving jobs through the request loop, two different
    subintegrations each, interleaved with several in flight: every result
    equals the same subintegration processed alone, bit for bit."""
    from test_gpu_serving import DeviceSubint, device_results, fetch
    lib, ctx = gpu
    host = fxcorr.load_host_lib()
    jobs, alone, subs = [], [], []
    for i, (nst, nchan) in enumerate(SERVING):
        per = []
        for k in range(2):
            spec, sd = serving_data(i, nst, nchan, k)
            if k == 0:
                jobs.append(spec.build())
            g = GpuJob(lib, ctx, jobs[-1][0])
            per.append((sd, g.process(sd)))
            g.close()
        alone.append(per)
        subs.append([DeviceSubint(lib, ctx, sd) for sd, _ in per])
    srv = C.c_void_p()
    check(host, host.fxs_create(0, 0, C.byref(srv)), "fxs_create")
    try:
        cids = []
        for job, keep in jobs:
            cid = C.c_int()
            check(host, host.fxs_add_config(srv, C.byref(job), 2, C.byref(cid)), "add_config")
            cids.append(cid.value)
        pending = []
        for rnd in range(2):
            for i in range(len(SERVING)):
                for k in range(2):
                    n = alone[i][k][1].size
                    res = device_results(lib, ctx, n)
                    req = C.c_int64()
                    ds = subs[i][k]
                    check(host, host.fxs_submit(srv, cids[i], C.byref(ds.sub), ds.st, res, C.byref(req)), "submit")
                    pending.append((req.value, i, k, res, n))
        for req, i, k, res, n in pending:
            check(host, host.fxs_wait(srv, req, None), "wait")
            # a second wait on a finished request is refused, not a use of freed memory
            assert host.fxs_wait(srv, req, None) == fxcorr.FXG_EINVAL
            got = fetch(lib, ctx, res, n)
            assert np.array_equal(got.view(np.uint32), alone[i][k][1].view(np.uint32)), (i, k)
            lib.fxg_device_free(ctx, res)
    finally:
        host.fxs_destroy(srv)
        for per in subs:
            for ds in per:
                ds.free()


# ---------------------------------------------------------------------------
# six uneven jobs in one graph launch
# ---------------------------------------------------------------------------
def test_uneven_jobs_batch_equals_alone_and_oracle(gpu, oracle):
    lib, ctx = gpu
    specs = uneven_workload_specs(xmac="fp16", max_blocks=3)
    jobs, datas, alone = [], [], []
    for k, (spec, label) in enumerate(specs):
        rng = np.random.default_rng(500 + k)
        sd = make_subint(spec, seed=500 + k, rho=0.05, delays_us=rng.uniform(-10.0, 10.0, spec.nstations),
                         rates_us_per_s=rng.uniform(-2.0, 2.0, spec.nstations), invalid_frames={1: [1]})
        job, keep = spec.build()
        g = GpuJob(lib, ctx, job)
        flush(f"uneven {label}: alone")
        alone.append(g.process(sd))
        jobs.append((job, keep, g))
        datas.append(sd)
    n = len(jobs)
    handles = (C.c_void_p * n)(*[g.h.value for _, _, g in jobs])
    b = C.c_void_p()
    check(lib, lib.fxg_batch_create(ctx, handles, n, C.byref(b)), "batch_create")
    try:
        subs = (fxcorr.SubintDesc * n)(*[sd.subint_desc() for sd in datas])
        sts = [sd.station_inputs() for sd in datas]
        starr = (C.POINTER(fxcorr.StationInput) * n)(*[C.cast(st, C.POINTER(fxcorr.StationInput)) for st, _ in sts])
        outs = [np.zeros(a.size, dtype=np.complex64) for a in alone]
        resarr = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
        for rep in range(2):                   # capture, then a replay
            for o in outs:
                o[:] = 0
            check(lib, lib.fxg_batch_process(b, subs, starr, 0, resarr, 0), "batch_process")
            for k in range(n):
                assert np.array_equal(outs[k].view(np.uint32), alone[k].view(np.uint32)), (rep, specs[k][1])
        caps = C.c_int64()
        check(lib, lib.fxg_batch_graph_captures(b, C.byref(caps)), "captures")
        assert caps.value == 1
    finally:
        lib.fxg_batch_destroy(b)
    for k, (spec, label) in enumerate(specs):
        job = jobs[k][0]
        flush(f"uneven {label}: FP64 oracle")
        ref = oracle_results(oracle, job, datas[k], fp64=True, nthreads=threads_for(spec))
        xend, a0, a1 = region_slices(oracle, job, plan_info(oracle, "orc_", job))
        ex, ea = rel_err(alone[k][:xend], ref[:xend]), rel_err(alone[k][a0:a1], ref[a0:a1])
        flush(f"  cross {ex:.3e} auto {ea:.3e}")
        assert ex < MIXED_TOL["fp16"] and ea < MIXED_TOL["fp16"], (label, ex, ea)
    for _, _, g in jobs:
        g.close()


# ---------------------------------------------------------------------------
# the DiFX writer on the library's results
# ---------------------------------------------------------------------------
def test_writer_records_from_gpu_results(gpu, oracle, tmp_path):
    from test_host_visibility import EXE, read_difx, write_job
    assert os.path.exists(EXE), "build first (__graft_entry__.build())"
    lib, ctx = gpu
    spec = JobSpec(nstations=4, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=24)
    job, keep = spec.build()
    rng = np.random.default_rng(61)
    sd = make_subint(spec, seed=61, rho=0.05, delays_us=rng.uniform(-20.0, 20.0, spec.nstations),
                     rates_us_per_s=rng.uniform(-1.0, 1.0, spec.nstations), invalid_frames={2: [1]})
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    g.close()
    ref = oracle_results(oracle, job, sd)
    recs = {}
    for name, res in (("gpu", got), ("oracle", ref)):
        d = tmp_path / name
        d.mkdir()
        write_job(str(d), job, sd)
        res.astype(np.complex64).tofile(str(d / "results.bin"))
        nsub, subintns = 3, 8_000_000
        for tsys in (0.0, 40.0):
            r = subprocess.run([EXE, str(d), str(nsub), str(tsys), str(subintns), str(nsub * subintns / 1e9)],
                               capture_output=True, text=True, timeout=60)
            assert r.returncode == 0, r.stdout + r.stderr
            recs[(name, tsys)] = read_difx(str(d / "out.difx"), job)
    for tsys in (0.0, 40.0):
        a, b = recs[("gpu", tsys)], recs[("oracle", tsys)]
        assert len(a) == len(b) > 0
        for ra, rb in zip(a, b):
            for key in ("baseline", "mjd", "seconds", "config", "source", "freq", "pol", "bin"):
                assert ra[key] == rb[key], key
            assert abs(ra["weight"] - rb["weight"]) <= 1e-6 * max(1.0, abs(rb["weight"]))
            assert rel_err(ra["data"], rb["data"]) < FP32_TOL, (ra["baseline"], ra["freq"], ra["pol"])


# ---------------------------------------------------------------------------
# Mode::process per block: a refilled buffer is prepared again
# ---------------------------------------------------------------------------
def test_process_block_sees_refilled_buffer(gpu, oracle):
    from fxtestlib import oracle_spectra
    lib, ctx = gpu
    spec = JobSpec(nstations=3, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, blockspersend=12)
    job, keep = spec.build()
    sd1 = make_subint(spec, seed=71, rho=0.1, delays_us=[0.0, 3.0, -7.5], rates_us_per_s=[0.0, 0.5, -0.3])
    sd2 = make_subint(spec, seed=72, rho=0.1, delays_us=[0.0, 3.0, -7.5], rates_us_per_s=[0.0, 0.5, -0.3])
    assert all(a.raw.size == b.raw.size for a, b in zip(sd1.stations, sd2.stations))
    bufs = [s.raw.copy() for s in sd1.stations]           # one host buffer per station, refilled below
    for s, buf in zip(sd1.stations, bufs):
        s.raw = buf
    g = GpuJob(lib, ctx, job)
    sub = sd1.subint_desc()
    st, k1 = sd1.station_inputs()
    nb = spec.nrecbands
    out = np.zeros(nb * spec.nchan, dtype=np.complex64)
    ds, blk = 1, 7
    check(lib, lib.fxg_station_process_block(g.h, C.byref(sub), st, 0, ds, blk, out.ctypes.data, None), "block")
    ref1 = oracle_spectra(oracle, job, sd1, ds, blk, nb, spec.nchan)
    assert rel_err(out, ref1.reshape(-1)) < FP32_TOL
    # same pointers, lengths, times and models; new bytes in the buffers
    for buf, s2 in zip(bufs, sd2.stations):
        buf[:] = s2.raw
    check(lib, lib.fxg_station_process_block(g.h, C.byref(sub), st, 0, ds, blk, out.ctypes.data, None), "block")
    ref2 = oracle_spectra(oracle, job, sd2, ds, blk, nb, spec.nchan)
    assert rel_err(out, ref2.reshape(-1)) < FP32_TOL
    assert rel_err(ref1, ref2) > 0.5
    g.close()
