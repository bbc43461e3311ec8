#!/usr/bin/env python3
"""Throughput benchmark of the MI355X FX correlator on the 32-station
mixed-precision workload.

Workload (one subintegration = one step): 32 stations x 2 pol x 16 subbands
x 64 MHz (128 Msamp/s real), 2-bit VDIF (8032-byte frames), 4096 channels
per subband (8192-point FFTs), 2000 FFT blocks (128 ms), every baseline with
RR/RL/LR/LL products plus auto and cross-pol autocorrelations.  Inputs are
seeded synthetic data (common Gaussian sky signal per subband, correlation
log-uniform in [1e-4, 1e-1], independent station noise) generated on the GPU,
so every timed step starts with its raw frames resident in HBM.

Prints one JSON line: input Gsamples/s of the mixed-precision (fp16 XMAC)
path, the plain FP32 path beside it, visibility error of every precision
against the FP64 oracle on a short sample of the same workload, the
roofline of the dominant kernel (live HIP-event timing) and the CPU
restatement timed on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "difx-svn2git-try3_amd"))

METRIC = "input Gsamples/s, mixed-precision vs plain FP32 at ≤1e-3 vis err; p99 latency"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp16": 2500.0, "bf16": 2500.0, "fp32": 157.3}
FP32_VECTOR_PEAK_TFLOPS = 256 * 64 * 2 * 2.4e9 / 1e12     # 78.6: one FMA per lane per clock, unpacked


# CPU threads of the baseline: the host share of one GPU on the benchmark
# machines (16), or fewer where the host has fewer cores
CPU_THREADS = max(1, min(16, os.cpu_count() or 1))
FP32_PEAK_TFLOPS = 157.3                # MI355X published FP32 vector peak (packed issue)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def profiled_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the 32-station workload
    (profiles/hbm_traffic_32st_4096ch.txt), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic_32st_4096ch.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_dispatch = "):
                return json.loads(line.split("=", 1)[1]).get(kernel)
    except OSError:
        return None
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def c3_spec(xmac: str, blocks: int, chunk: int = 0):
    from fxcorr import JobSpec
    return JobSpec(nstations=32, nsubbands=16, npol=2, bw_mhz=64.0, nchan=4096, nbits=2,
                   blockspersend=blocks, crosspol=True, xmac=xmac, framebytes=8032,
                   base_freq_mhz=8200.0, max_chunk_blocks=chunk)


def rhos_for(nsb, seed):
    rng = np.random.default_rng(seed)
    return 10.0 ** rng.uniform(-4.0, -1.0, nsb)


class Job:
    def __init__(self, lib, ctx, spec):
        from fxcorr import PlanInfo, check
        self.lib, self.spec = lib, spec
        self.job, self.keep = spec.build()
        self.h = C.c_void_p()
        check(lib, lib.fxg_job_create(ctx, C.byref(self.job), C.byref(self.h)), "fxg_job_create")
        self.info = PlanInfo()
        check(lib, lib.fxg_job_plan_info(self.h, C.byref(self.info)), "plan_info")

    def run(self, sub, st, dres):
        from fxcorr import check
        check(self.lib, self.lib.fxg_subint_process(self.h, C.byref(sub), st, 1, C.c_void_p(dres), 1),
              "fxg_subint_process")

    def stage_times(self, reset=True):
        from fxcorr import STAGE_NAMES
        ms = (C.c_double * len(STAGE_NAMES))()
        n = (C.c_int64 * len(STAGE_NAMES))()
        self.lib.fxg_job_stage_times(self.h, ms, n, 1 if reset else 0)
        return list(ms), list(n)

    def close(self):
        self.lib.fxg_job_destroy(self.h)


def time_job(lib, ctx, spec, sd, raws, steps, warmup, torch, autotune=False, graph=False):
    from fxsynth import station_inputs_device
    from fxcorr import Schedule, TuneResult, check
    job = Job(lib, ctx, spec)
    sub = sd.subint_desc()
    st, keep = station_inputs_device(sd, raws)
    dres = torch.empty(2 * job.info.result_length, dtype=torch.float32, device="cuda:0")
    tuning = None
    if autotune:
        res = (TuneResult * 32)()
        nres = C.c_int()
        check(lib, lib.fxg_job_autotune(job.h, C.byref(sub), st, 1, 1, res, 32, C.byref(nres)), "autotune")
        tuning = [{"schedule": res[i].schedule.as_dict(), "ms": res[i].ms, "rel_err": res[i].rel_err,
                   "accepted": bool(res[i].accepted)} for i in range(min(nres.value, 32))]
        log(f"  autotune: {len(tuning)} schedules tried")
    sched = Schedule()
    check(lib, lib.fxg_job_get_schedule(job.h, C.byref(sched)), "get_schedule")
    if graph:
        check(lib, lib.fxg_job_set_graph(job.h, 1), "set_graph")
    for _ in range(warmup):
        job.run(sub, st, dres.data_ptr())
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        job.run(sub, st, dres.data_ptr())      # synchronous: returns when the subint is done
        lat.append(time.perf_counter() - a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # one extra profiled step for the per-stage device times
    lib.fxg_job_set_profiling(job.h, 1)
    job.stage_times(reset=True)
    job.run(sub, st, dres.data_ptr())
    ms, n = job.stage_times(reset=True)
    lib.fxg_job_set_profiling(job.h, 0)
    info = job.info
    job.close()
    del dres
    time_job.schedule = {"schedule": sched.as_dict(), "tuning": tuning, "graph": graph}
    time_job.sched = sched
    return el, lat, ms, n, info


def pcie_inclusive(lib, ctx, spec, sd, raws, steps, torch, sched, graph):
    """The same subintegration with the raw VDIF handed over in pinned host
    memory.  Submitted back to back (fxg_subint_submit), each subint's PCIe
    copy runs on the job's copy stream under the previous one's kernels;
    one at a time (fxg_subint_process), copy and kernels add up.  Reported
    beside the headline, never as it."""
    from fxcorr import STAGE_NAMES, check
    from fxsynth import station_inputs_device
    host = []
    for r in raws:
        h = torch.empty(r.numel(), dtype=torch.uint8, pin_memory=True)
        h.copy_(r)
        host.append(h)
    torch.cuda.synchronize()
    job = Job(lib, ctx, spec)
    check(lib, lib.fxg_job_set_schedule(job.h, C.byref(sched)), "set_schedule")
    if graph:
        check(lib, lib.fxg_job_set_graph(job.h, 1), "set_graph")
    sub = sd.subint_desc()
    st, keep = station_inputs_device(sd, host)           # host pointers this time
    dres = torch.empty(2 * job.info.result_length, dtype=torch.float32, device="cuda:0")

    def run():
        check(lib, lib.fxg_subint_process(job.h, C.byref(sub), st, 0, C.c_void_p(dres.data_ptr()), 1),
              "fxg_subint_process")

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    el_sync = time.perf_counter() - t0
    # nsub back-to-back submissions, timed from the first submit to the last
    # completion: the pipeline's fill (the first copy, with no kernels under
    # it) and drain (the last kernels) included
    nsub = 6 * steps
    t0 = time.perf_counter()
    ticket = C.c_int64()
    for _ in range(nsub):
        check(lib, lib.fxg_subint_submit(job.h, C.byref(sub), st, 0, C.c_void_p(dres.data_ptr()), 1,
                                         C.byref(ticket)), "fxg_subint_submit")
    check(lib, lib.fxg_subint_wait(job.h, ticket.value), "fxg_subint_wait")
    el = (time.perf_counter() - t0) / nsub * steps
    lib.fxg_job_set_profiling(job.h, 1)
    job.stage_times(reset=True)
    run()
    ms, n = job.stage_times(reset=True)
    lib.fxg_job_set_profiling(job.h, 0)
    job.close()
    nbytes = sum(int(h.numel()) for h in host)
    del dres, host
    h2d = ms[STAGE_NAMES.index("h2d")]
    samples = spec.nstations * spec.nrecbands * spec.blockspersend * spec.fftchannels
    return {"value": samples * steps / el / 1e9, "unit": "Gsamples/s", "ms_per_step": el / steps * 1e3,
            "submissions": nsub, "one_at_a_time": {"value": samples * steps / el_sync / 1e9,
                                                  "ms_per_step": el_sync / steps * 1e3, "steps": steps},
            "h2d_ms": h2d, "h2d_gbs": nbytes / (h2d * 1e-3) / 1e9 if h2d > 0 else None,
            "raw_bytes": nbytes,
            "input": "raw VDIF in pinned host memory; submitted back to back, each subint's copy overlaps "
                     "the previous subint's kernels; value over all submissions, first submit to last "
                     "completion"}


def vis_errors(lib, ctx, orc, seed, blocks, threads, torch):
    """Visibility error of every XMAC precision against the FP64 oracle on a
    short (blocks-long) subintegration of the same workload."""
    from fxsynth import make_subint_device, station_inputs_device
    from fxcorr import check
    spec64 = c3_spec("fp32", blocks)
    rng = np.random.default_rng(seed)
    delays = [int(v) for v in rng.integers(-1500, 1500, spec64.nstations)]
    sd, raws = make_subint_device(spec64, seed=seed, rho=rhos_for(spec64.nsubbands, seed), delay_samples=delays)
    host = [r.cpu().numpy() for r in raws]
    for s, h in zip(sd.stations, host):
        s.raw = h
    job64, keep = spec64.build()
    from fxcorr import PlanInfo
    info = PlanInfo()
    check(orc, orc.orc_plan(C.byref(job64), C.byref(info)), "orc_plan")
    ref = np.zeros(2 * info.result_length, dtype=np.float64)
    sub = sd.subint_desc()
    st, k = sd.station_inputs()
    t = time.perf_counter()
    check(orc, orc.orc_process(C.byref(job64), C.byref(sub), st, 1, threads, 0, blocks, ref.ctypes.data), "orc64")
    log(f"fp64 oracle on {blocks} blocks: {time.perf_counter() - t:.1f}s")
    ref = ref.view(np.complex128)
    off = C.c_int64()
    check(orc, orc.orc_plan_bweight_offset(C.byref(job64), 0, 0, C.byref(off)), "bw")
    xend = off.value
    check(orc, orc.orc_plan_autocorr_offset(C.byref(job64), 0, C.byref(off)), "ac")
    a0 = off.value
    # per-(datastream, band) parallel-hand autocorrelation spectra of the oracle
    nb, nch = spec64.nrecbands, spec64.nchan
    acs = []
    for d in range(spec64.nstations):
        check(orc, orc.orc_plan_autocorr_offset(C.byref(job64), d, C.byref(off)), "ac")
        acs.append(ref[off.value: off.value + nb * nch].real.reshape(nb, nch))
    out = {}
    for prec in ("fp16", "bf16", "fp8", "fp32"):
        spec = c3_spec(prec, blocks)
        job = Job(lib, ctx, spec)
        dres = torch.empty(2 * job.info.result_length, dtype=torch.float32, device="cuda:0")
        sti, keep2 = station_inputs_device(sd, raws)
        job.run(sub, sti, dres.data_ptr())
        if prec == "fp8":
            # per-channel power-of-two scales from this subintegration's power, then again
            check(lib, lib.fxg_job_calibrate_scales(job.h), "calibrate")
            job.run(sub, sti, dres.data_ptr())
        got = dres.cpu().numpy().view(np.complex64).astype(np.complex128)
        job.close()
        d = got[:xend] - ref[:xend]
        rel = float(np.linalg.norm(d) / np.linalg.norm(ref[:xend]))
        # correlation-coefficient units: |dV| / sqrt(A1 A2) per channel
        coef = 0.0
        bl = 0
        pos = 0
        for f in range(spec.nsubbands):
            for i in range(spec.nstations):
                for j in range(i + 1, spec.nstations):
                    for (p1, p2) in ((0, 0), (0, 1), (1, 0), (1, 1)):
                        a1 = acs[i][f * 2 + p1]
                        a2 = acs[j][f * 2 + p2]
                        seg = d[pos: pos + nch]
                        coef = max(coef, float(np.max(np.abs(seg) / np.sqrt(a1 * a2))))
                        pos += nch
        out[prec] = {"rel": rel, "coef_max": coef}
        del dres
    return out


def roofline(spec, info, stage, launch, precision):
    """Roofline of the step's dominant kernel from the profiled step's HIP
    event times: algorithmic bytes and flops per launch against HBM and the
    FP32 vector (F stage) or matrix-core (XMAC) peak; "bound" names the
    resource with the larger fraction."""
    C_, L = spec.nchan, spec.fftchannels
    nin = spec.nstations * spec.npol
    specbytes = 8 if precision == "fp32" else 4
    dom = max(("fstage", "xmac"), key=lambda k: stage.get(k, {"ms": 0})["ms"])
    nblk = info.chunk_blocks
    samples = spec.nstations * spec.nrecbands * nblk * L
    if dom == "fstage":
        avg_s = stage["fstage"]["ms"] / stage["fstage"]["launches"] / 1e3
        kname = f"fxg::k_fstage<{L}, 2>"
        # 2-bit input read once + spectra written once
        nbytes = samples * 0.25 + spec.nstations * spec.nrecbands * nblk * C_ * specbytes
        # FFT 5 log2 L, fringe rotation 6, fractional-sample rotation 3 flops per sample
        flops = samples * (5.0 * np.log2(L) + 9.0)
        hbm_frac = nbytes / avg_s / 1e9 / HBM_PEAK_GBS
        valu_tf = flops / avg_s / 1e12
        roof = {"kernel": kname, "avg_launch_ms": avg_s * 1e3, "traffic": profiled_traffic(kname),
                "algorithmic_bytes_per_launch": nbytes, "flops_per_launch": flops,
                "hbm_gbs": nbytes / avg_s / 1e9, "hbm_frac": hbm_frac,
                "fp32_tflops": valu_tf, "fp32_frac": valu_tf / FP32_PEAK_TFLOPS,
                "fp32_unpacked_peak_tflops": FP32_VECTOR_PEAK_TFLOPS,
                "fp32_unpacked_frac": valu_tf / FP32_VECTOR_PEAK_TFLOPS}
        if valu_tf / FP32_PEAK_TFLOPS >= hbm_frac:
            roof.update(bound="valu", achieved=valu_tf, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                        frac=valu_tf / FP32_PEAK_TFLOPS)
        else:
            roof.update(bound="hbm", achieved=nbytes / avg_s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s", frac=hbm_frac)
        return roof
    # XMAC: Hermitian pairs incl. autos, 8 flops per complex MAC; spectra read once
    avg_s = stage["xmac"]["ms"] / stage["xmac"]["launches"] / 1e3
    nfreq = stage["xmac"]["launches"]
    flops = 8.0 * nin * (nin + 1) / 2 * C_ * nblk * spec.nsubbands / nfreq
    nbytes = nin * C_ * nblk * specbytes * spec.nsubbands / nfreq
    peak = PEAK_TFLOPS[precision]
    bf = "true" if precision == "bf16" else "false"
    kname = (f"fxg::k_xmac_mfma_dl<2, {bf}, false>" if launch["schedule"]["xmac_streaming"]
             else f"fxg::k_xmac_mfma<2, {bf}, 16>")
    tf = flops / avg_s / 1e12
    hbm_frac = nbytes / avg_s / 1e9 / HBM_PEAK_GBS
    roof = {"kernel": kname, "avg_launch_ms": avg_s * 1e3, "traffic": profiled_traffic(kname),
            "flops_per_launch": flops, "algorithmic_bytes_per_launch": nbytes, "tflops": tf, "mfma_frac": tf / peak,
            "hbm_frac": hbm_frac}
    if tf / peak >= hbm_frac:
        roof.update(bound="mfma", achieved=tf, peak=peak, unit="TFLOP/s", frac=tf / peak)
    else:
        roof.update(bound="hbm", achieved=nbytes / avg_s / 1e9, peak=HBM_PEAK_GBS, unit="GB/s", frac=hbm_frac)
    return roof


def small_config_cpu_baselines(orc, lib, ctx, torch):
    """The restated reference CPU path on the two small configurations:
    4 stations x 2 pol x 16 MHz, 256 channels, eight 1 s subintegrations in
    full; and the first 4096 FFT blocks (131 ms) of a 1 s subintegration of
    10 stations x 2 pol x 8 x 32 MHz, 1024 channels.  CPU_THREADS threads,
    each a contiguous block range (core.cpp:2091-2095)."""
    from fxcorr import JobSpec
    from fxsynth import make_subint_device
    out = {}
    rng = np.random.default_rng(11)
    spec = JobSpec(nstations=4, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, blockspersend=62500, crosspol=True,
                   xmac="fp32")
    tot_el, tot_ns = 0.0, 0
    for k in range(8):
        delays = [int(v) for v in rng.uniform(-50.0, 50.0, spec.nstations) * 32]     # us -> samples at 32 Ms/s
        sd, raws = make_subint_device(spec, seed=100 + k, rho=0.05, delay_samples=delays,
                                      rates_us_per_s=rng.uniform(-1.0, 1.0, spec.nstations), sub_sec=10 + k)
        for s_, r in zip(sd.stations, raws):
            s_.raw = r.cpu().numpy()
        del raws
        v, el, ns = cpu_baseline(orc, sd, spec, CPU_THREADS, spec.blockspersend)
        tot_el += el
        tot_ns += ns
    out["4st_256ch_8s"] = {"value": tot_ns / tot_el / 1e9, "unit": "Gsamples/s", "cores": CPU_THREADS,
                           "wall_s": tot_el, "gsamples": tot_ns / 1e9,
                           "sample": "eight 1 s subintegrations (62500 blocks of 512 each), in full"}
    log(f"  CPU 4-station 8 s: {out['4st_256ch_8s']['value']:.4f} Gsamples/s in {tot_el:.1f} s")
    spec = JobSpec(nstations=10, nsubbands=8, npol=2, bw_mhz=32.0, nchan=1024, blockspersend=4096, crosspol=True,
                   xmac="fp32")
    delays = [int(v) for v in rng.uniform(-10000.0, 10000.0, spec.nstations) * 64]    # +-10 ms at 64 Ms/s
    sd, raws = make_subint_device(spec, seed=200, rho=0.02, delay_samples=delays,
                                  rates_us_per_s=rng.uniform(-3.0, 3.0, spec.nstations))
    for s_, r in zip(sd.stations, raws):
        s_.raw = r.cpu().numpy()
    del raws
    v, el, ns = cpu_baseline(orc, sd, spec, CPU_THREADS, spec.blockspersend)
    out["10st_1024ch_slice"] = {"value": v, "unit": "Gsamples/s", "cores": CPU_THREADS, "wall_s": el,
                                "gsamples": ns / 1e9,
                                "sample": "first 4096 FFT blocks of 2048 (131 ms) of a 1 s subintegration"}
    log(f"  CPU 10-station slice: {v:.4f} Gsamples/s in {el:.1f} s")
    out["cpu_model"] = cpu_model()
    out["nproc"] = os.cpu_count()
    torch.cuda.empty_cache()
    return out


def cpu_baseline(orc, sd_host, spec, threads, nblocks):
    from fxcorr import PlanInfo, check
    job, keep = spec.build()
    info = PlanInfo()
    check(orc, orc.orc_plan(C.byref(job), C.byref(info)), "orc_plan")
    out = np.zeros(2 * info.result_length, dtype=np.float32)
    sub = sd_host.subint_desc()
    st, k = sd_host.station_inputs()
    t = time.perf_counter()
    check(orc, orc.orc_process(C.byref(job), C.byref(sub), st, 0, threads, 0, nblocks, out.ctypes.data), "orc")
    el = time.perf_counter() - t
    samples = spec.nstations * spec.nrecbands * nblocks * spec.fftchannels
    return samples / el / 1e9, el, samples


# The serving workload: three concurrent correlation jobs (stations,
# channels) of 8 subbands x 64 MHz, 2-bit, streaming 0.125 s subints.
SERVING_JOBS = [(10, 1024), (20, 4096), (32, 8192)]
SERVING_SUBINT_S = 0.125


def serving_bench(args):
    """Poisson arrivals of 0.125 s subintegrations of three jobs at a multiple
    of the real-time aggregate rate, through the request loop
    (include/fxcorr_serve.h); p50/p99 latency (completion - submission) and
    sustained input Gsamples/s."""
    import threading
    import torch
    import fxcorr
    from fxcorr import JobSpec, ServeStats, check
    from fxsynth import make_subint_device, station_inputs_device
    lib = fxcorr.load_gpu_lib()
    host = fxcorr.load_host_lib()
    srv = C.c_void_p()
    check(host, host.fxs_create(0, 0, C.byref(srv)), "fxs_create")
    check(host, host.fxs_set_max_inflight(srv, args.inflight), "fxs_set_max_inflight")
    cfgs = []
    for i, (nst, nchan) in enumerate(SERVING_JOBS):
        L = 2 * nchan
        blocks = int(round(SERVING_SUBINT_S * 128e6 / L))
        spec = JobSpec(nstations=nst, nsubbands=8, npol=2, bw_mhz=64.0, nchan=nchan, blockspersend=blocks,
                       xmac=args.precision)
        job, keep = spec.build()
        cid = C.c_int()
        check(host, host.fxs_add_config(srv, C.byref(job), args.slots, C.byref(cid)), "fxs_add_config")
        sd, raws = make_subint_device(spec, seed=args.seed + i, rho=rhos_for(8, args.seed + i))
        st, kst = station_inputs_device(sd, raws)
        info = fxcorr.PlanInfo()
        check(lib, lib.fxg_plan(C.byref(job), C.byref(info)), "plan")
        ring = [torch.empty(2 * info.result_length, dtype=torch.float32, device="cuda:0") for _ in range(args.slots + 2)]
        samples = nst * 16 * blocks * L
        cfgs.append(dict(cid=cid.value, job=job, keep=keep, sd=sd, raws=raws, st=st, kst=kst, sub=sd.subint_desc(),
                         ring=ring, samples=samples, nst=nst, nchan=nchan, blocks=blocks))
        log(f"job {i}: {nst} stations, {nchan} ch, {blocks} blocks/subint, {samples / 1e9:.2f} Gsamples/subint")
    torch.cuda.synchronize()
    rate = args.load / SERVING_SUBINT_S            # arrivals per second per job
    rng = np.random.default_rng(args.seed)
    events = []
    for i in range(len(cfgs)):
        t = 0.0
        while True:
            t += rng.exponential(1.0 / rate)
            if t > args.duration:
                break
            events.append((t, i))
    events.sort()
    log(f"{len(events)} subints over {args.duration} s at {args.load}x real time")
    # warm every configuration's first slot before the clock starts
    for c in cfgs:
        req = C.c_int64()
        check(host, host.fxs_submit(srv, c["cid"], C.byref(c["sub"]), c["st"], C.c_void_p(c["ring"][0].data_ptr()),
                                    C.byref(req)), "warm")
        check(host, host.fxs_wait(srv, req, None), "warm wait")
    records = []                                   # (job, request, ring slot)
    lock = threading.Lock()
    inflight = [dict() for _ in cfgs]              # ring index -> request of the job
    done = []
    cv = threading.Condition(lock)

    def waiter():
        while True:
            with cv:
                while not records:
                    cv.wait()
                item = records.pop(0)
            if item is None:
                return
            i, req, k, t_sub = item
            lat = C.c_int64()
            check(host, host.fxs_wait(srv, req, C.byref(lat)), "fxs_wait")
            with cv:
                done.append((i, lat.value / 1e6, time.perf_counter()))
                inflight[i].pop(k, None)
                cv.notify_all()

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    t0 = time.perf_counter()
    nxt = [0] * len(cfgs)
    for t_arr, i in events:
        now = time.perf_counter() - t0
        if t_arr > now:
            time.sleep(t_arr - now)
        c = cfgs[i]
        with cv:                                   # a result buffer of this job that is free
            k = nxt[i] % len(c["ring"])
            while k in inflight[i]:
                cv.wait()
            nxt[i] += 1
        req = C.c_int64()
        check(host, host.fxs_submit(srv, c["cid"], C.byref(c["sub"]), c["st"], C.c_void_p(c["ring"][k].data_ptr()),
                                    C.byref(req)), "fxs_submit")
        with cv:
            inflight[i][k] = req.value
            records.append((i, req.value, k, time.perf_counter()))
            cv.notify_all()
    with cv:
        records.append(None)
        cv.notify_all()
    th.join()
    t_end = max(d[2] for d in done)
    el = t_end - t0
    stats = ServeStats()
    check(host, host.fxs_stats(srv, C.byref(stats)), "stats")
    lat = np.array([d[1] for d in done])
    per = {}
    for i, c in enumerate(cfgs):
        li = np.array([d[1] for d in done if d[0] == i])
        per[f"{c['nst']}st_{c['nchan']}ch"] = {"subints": int(li.size), "p50_ms": float(np.percentile(li, 50)),
                                             "p99_ms": float(np.percentile(li, 99))}
    samples = sum(cfgs[d[0]]["samples"] for d in done)
    offered = sum(c["samples"] for c in cfgs) / SERVING_SUBINT_S * args.load / 1e9
    out = {
        "metric": "serving: sustained input Gsamples/s and subint latency, 3 concurrent jobs",
        "value": samples / el / 1e9, "unit": "Gsamples/s", "n_gpus": 1, "higher_is_better": True,
        "offered_gsamples_per_s": offered, "load_vs_realtime": args.load, "duration_s": el,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "per_job": per, "dtype": args.precision, "data": "synthetic 2-bit VDIF generated on the GPU",
        "config": {"workload": "3 jobs (10/20/32 stations; 1024/4096/8192 ch) x 8 subbands x 64 MHz, 2-bit, "
                               "0.125 s subints, Poisson arrivals", "slots_per_job": args.slots,
                   "max_inflight": args.inflight},
        "server": {"budget_gb": stats.budget_bytes / 1e9, "used_gb": stats.used_bytes / 1e9,
                   "slots": stats.nslots, "refused_slots": stats.refused},
    }
    host.fxs_destroy(srv)
    return out


def uneven_bench(args):
    """Six uneven jobs (4..30 stations, 64..32768 channels, 2/8-bit, 4..16
    subbands of 16..128 MHz, two zoom bands each) of one 0.128 s
    subintegration each, all in one graph launch per round (fxg_batch), data
    resident in HBM; also each job alone, one after another."""
    import torch
    import fxcorr
    from fxcorr import check
    from fxsynth import make_subint_device, station_inputs_device, uneven_workload_specs

    lib = fxcorr.load_gpu_lib()
    ctx = C.c_void_p()
    check(lib, lib.fxg_ctx_create(0, C.byref(ctx)), "fxg_ctx_create")
    specs = uneven_workload_specs(args.subint, xmac=args.precision)
    rng = np.random.default_rng(args.seed)
    jobs, keep, subs, sts, dres, samples, labels = [], [], [], [], [], [], []
    for k, (spec, label) in enumerate(specs):
        log(f"job {label}: {spec.blockspersend} blocks of {spec.fftchannels}")
        delays = [int(v) for v in rng.integers(-1500, 1500, spec.nstations)]
        sd, raws = make_subint_device(spec, seed=args.seed + k, rho=rhos_for(spec.nsubbands, args.seed + k),
                                      delay_samples=delays)
        j = Job(lib, ctx, spec)
        st, kst = station_inputs_device(sd, raws)
        out = torch.empty(2 * j.info.result_length, dtype=torch.float32, device="cuda:0")
        if args.autotune:
            res = (fxcorr.TuneResult * 32)()
            n = C.c_int()
            sub = sd.subint_desc()
            check(lib, lib.fxg_job_autotune(j.h, C.byref(sub), st, 1, 1, res, 32, C.byref(n)), "autotune")
        if args.graph:
            check(lib, lib.fxg_job_set_graph(j.h, 1), "set_graph")
        jobs.append(j)
        keep.append((sd, raws, st, kst, out))
        subs.append(sd.subint_desc())
        sts.append(st)
        dres.append(out)
        samples.append(spec.nstations * spec.nrecbands * spec.blockspersend * spec.fftchannels)
        labels.append(label)
    torch.cuda.synchronize()
    total = sum(samples)

    # each job alone, one after another
    alone = []
    for j, sub, st, out in zip(jobs, subs, sts, dres):
        j.run(sub, st, out.data_ptr())
        torch.cuda.synchronize()
        t = time.perf_counter()
        j.run(sub, st, out.data_ptr())
        torch.cuda.synchronize()
        alone.append(time.perf_counter() - t)
    seq_value = total / sum(alone) / 1e9

    # all six in one graph launch per round
    n = len(jobs)
    arr = (C.c_void_p * n)(*[j.h.value for j in jobs])
    b = C.c_void_p()
    check(lib, lib.fxg_batch_create(ctx, arr, n, C.byref(b)), "batch_create")
    subarr = (fxcorr.SubintDesc * n)(*subs)
    starr = (C.POINTER(fxcorr.StationInput) * n)(*[C.cast(st, C.POINTER(fxcorr.StationInput)) for st in sts])
    resarr = (C.c_void_p * n)(*[o.data_ptr() for o in dres])
    for _ in range(args.warmup):
        check(lib, lib.fxg_batch_process(b, subarr, starr, 1, resarr, 1), "batch_process")
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        check(lib, lib.fxg_batch_process(b, subarr, starr, 1, resarr, 1), "batch_process")
        lat.append(time.perf_counter() - a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = total * args.steps / el / 1e9
    lat_ms = np.array(lat) * 1e3
    log(f"  batch {value:.1f} Gsamples/s ({el / args.steps * 1e3:.1f} ms/round); one by one {seq_value:.1f}")
    lib.fxg_batch_destroy(b)
    out = {
        "metric": "input Gsamples/s, six uneven jobs in one graph launch",
        "value": value, "unit": "Gsamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.precision, "data": "synthetic, resident in HBM",
        "config": {"workload": "6 uneven jobs, one subint of each per step",
                   "jobs": labels, "subint_s": args.subint, "samples_per_step": total},
        "one_by_one": {"value": seq_value, "ms": [a * 1e3 for a in alone]},
        "realtime_rate": total / args.subint / 1e9,
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
    }
    for j in jobs:
        j.close()
    lib.fxg_ctx_destroy(ctx)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--blocks", type=int, default=2000, help="FFT blocks per subintegration")
    ap.add_argument("--precision", default="fp16", choices=["fp16", "bf16", "fp8", "fp32"])
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 comparison")
    ap.add_argument("--no-verify", action="store_true", help="skip the FP64-oracle error check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-blocks", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=2000,
                    help="FFT blocks of spectra resident in HBM at once (default: the whole subint)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--serving", action="store_true",
                    help="the request-loop workload instead: 3 concurrent jobs, Poisson subint arrivals")
    ap.add_argument("--load", type=float, default=1.5, help="serving: offered load as a multiple of real time")
    ap.add_argument("--duration", type=float, default=8.0, help="serving: seconds of arrivals")
    ap.add_argument("--slots", type=int, default=2, help="serving: concurrent slots per job")
    ap.add_argument("--inflight", type=int, default=2,
                    help="serving: requests on the device at once, started first come first served (0: no bound)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the library's autotuner (XMAC schedule and pass length)")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="launch the kernels one by one instead of one graph per subintegration")
    ap.add_argument("--uneven", action="store_true",
                    help="the six-job uneven workload in one graph launch per step instead of the 32-station job")
    ap.add_argument("--subint", type=float, default=0.128, help="uneven: subintegration length in seconds")
    ap.add_argument("--no-serving", action="store_true", help="default run: skip the serving workload")
    ap.add_argument("--no-uneven", action="store_true", help="default run: skip the six uneven jobs")
    ap.add_argument("--no-pcie", action="store_true",
                    help="skip the run with the raw data in pinned host memory (PCIe copy included)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()

    if args.serving:
        if args.gpus != 1:
            print("bench.py runs on one GPU; --gpus above 1 is not supported", file=sys.stderr)
            sys.exit(2)
        print(json.dumps(serving_bench(args)), flush=True)
        return

    if args.uneven:
        if args.gpus != 1:
            print("bench.py runs on one GPU; --gpus above 1 is not supported", file=sys.stderr)
            sys.exit(2)
        print(json.dumps(uneven_bench(args)), flush=True)
        return

    if args.copy_calibration:
        import torch
        n = 1 << 30
        src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(7)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 3}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import fxcorr
    from fxcorr import check
    from fxsynth import make_subint_device

    lib = fxcorr.load_gpu_lib()
    ctx = C.c_void_p()
    check(lib, lib.fxg_ctx_create(0, C.byref(ctx)), "fxg_ctx_create")

    spec = c3_spec(args.precision, args.blocks, args.chunk)
    rng = np.random.default_rng(args.seed)
    delays = [int(v) for v in rng.integers(-1500, 1500, spec.nstations)]
    log("generating synthetic subintegration on the GPU ...")
    t = time.perf_counter()
    sd, raws = make_subint_device(spec, seed=args.seed, rho=rhos_for(spec.nsubbands, args.seed),
                                  delay_samples=delays)
    torch.cuda.synchronize()
    log(f"  {sum(r.numel() for r in raws) / 1e9:.2f} GB of VDIF in {time.perf_counter() - t:.1f}s")
    samples = spec.nstations * spec.nrecbands * spec.blockspersend * spec.fftchannels

    log(f"timing {args.precision}: {args.warmup} warmup + {args.steps} steps")
    el, lat, ms, n, info = time_job(lib, ctx, spec, sd, raws, args.steps, args.warmup, torch,
                                    autotune=args.autotune, graph=args.graph)
    launch = time_job.schedule
    value = samples * args.steps / el / 1e9
    log(f"  {value:.1f} Gsamples/s  ({el / args.steps * 1e3:.1f} ms/subint)")
    stage = {fxcorr.STAGE_NAMES[i]: {"ms": ms[i], "launches": n[i]} for i in range(len(ms)) if n[i]}

    pcie = None
    if not args.no_pcie:
        log("timing with the raw data in pinned host memory (PCIe copy included)")
        pcie = pcie_inclusive(lib, ctx, spec, sd, raws, 2, torch, time_job.sched, args.graph)
        log(f"  {pcie['value']:.1f} Gsamples/s ({pcie['ms_per_step']:.1f} ms/subint back to back; one at a "
            f"time {pcie['one_at_a_time']['ms_per_step']:.1f} ms, h2d {pcie['h2d_ms']:.1f} ms)")

    fp32 = None
    if not args.no_fp32 and args.precision != "fp32":
        log("timing plain fp32")
        spec32 = c3_spec("fp32", args.blocks, args.chunk)
        el32, lat32, ms32, n32, info32 = time_job(lib, ctx, spec32, sd, raws, max(1, args.steps // 2), 1, torch,
                                                  autotune=args.autotune, graph=args.graph)
        launch32 = time_job.schedule
        v32 = samples * max(1, args.steps // 2) / el32 / 1e9
        fp32 = {"value": v32, "ms_per_step": el32 / max(1, args.steps // 2) * 1e3, "launch": launch32,
                "stages_ms": {fxcorr.STAGE_NAMES[i]: ms32[i] for i in range(len(ms32)) if n32[i]}}
        log(f"  fp32 {v32:.1f} Gsamples/s")

    # ---- roofline of the dominant kernel (from the profiled step)
    roof = roofline(spec, info, stage, launch, args.precision)

    errors = None
    if not args.no_verify:
        log("visibility error vs the FP64 oracle on a 6-block sample")
        orc = fxcorr.load_oracle()
        errors = vis_errors(lib, ctx, orc, args.seed + 1, 6, CPU_THREADS, torch)
        log(f"  {errors}")

    cpu = cpu_extra = None
    if not args.no_cpu:
        orc = fxcorr.load_oracle()
        nb = args.cpu_blocks or 4 * CPU_THREADS
        host = [r.cpu().numpy() for r in raws]
        for s_, h in zip(sd.stations, host):
            s_.raw = h
        spec_cpu = c3_spec("fp32", args.blocks)
        log(f"CPU baseline: oracle FP32, {CPU_THREADS} threads, {nb} blocks of the 32-station workload")
        v, el_cpu, ns = cpu_baseline(orc, sd, spec_cpu, CPU_THREADS, nb)
        cpu = {"value": v, "unit": "Gsamples/s", "cores": CPU_THREADS, "kind": "port",
               "sample": f"first {nb} of {args.blocks} FFT blocks of the same subintegration "
                         f"({ns / 1e9:.3f} Gsamples, {el_cpu:.1f} s wall), {CPU_THREADS} threads each "
                         f"taking a contiguous block range (core.cpp:2091-2095)",
               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        log(f"  {v:.4f} Gsamples/s")
        del host
        for s_ in sd.stations:
            s_.raw = None
        cpu_extra = small_config_cpu_baselines(orc, lib, ctx, torch)

    lat_ms = np.array(lat) * 1e3
    job_lat = {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))}
    del raws, sd
    torch.cuda.empty_cache()
    serving = uneven = None
    if not args.no_serving:
        log("serving workload (3 concurrent jobs, Poisson arrivals)")
        sargs = argparse.Namespace(**vars(args))
        sargs.duration = min(args.duration, 6.0)
        serving = serving_bench(sargs)
        serving = {k: serving[k] for k in ("value", "unit", "offered_gsamples_per_s", "load_vs_realtime",
                                          "duration_s", "latency_ms", "per_job", "server", "config")}
        torch.cuda.empty_cache()
    if not args.no_uneven:
        log("six uneven jobs in one graph launch")
        uargs = argparse.Namespace(**vars(args))
        uargs.steps, uargs.warmup = 3, 1
        uneven = uneven_bench(uargs)
        uneven = {k: uneven[k] for k in ("value", "unit", "ms_per_step", "one_by_one", "realtime_rate",
                                         "latency_ms", "config")}
        torch.cuda.empty_cache()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded Gaussian sky + station noise, 2-bit VDIF generated on the GPU)",
        "config": {"workload": "32 stations x 2 pol x 16 subbands x 64 MHz, 2-bit VDIF, 4096 ch, "
                               f"{args.blocks} FFT blocks of 8192 per subint ({args.blocks * 8192 / 128e6 * 1e3:.0f} ms)",
                   "stations": spec.nstations, "subbands": spec.nsubbands, "pols": spec.npol,
                   "channels": spec.nchan, "blocks_per_subint": spec.blockspersend,
                   "xmac": args.precision, "fft": "fp32", "launch": launch},
        "latency_ms": job_lat,
        "stages_ms": {k: v["ms"] for k, v in stage.items()},
        "plain_fp32": fp32,
        "pcie_inclusive": pcie,
        "mixed_vs_fp32": (value / fp32["value"]) if fp32 else None,
        "vis_error_vs_fp64": errors,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_small_configs": cpu_extra,
        "serving": serving,
        "uneven_6jobs": uneven,
    }
    lib.fxg_ctx_destroy(ctx)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
