"""Graph-captured subintegrations and batches of jobs in one graph launch:
every replay equals the kernels launched one by one (the same kernels, so
bit for bit), with eager work and host synchronisations between replays,
and a change of launch shape captures a new graph."""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import JobSpec, check
from fxsynth import make_subint
from fxtestlib import GpuJob

pytestmark = pytest.mark.gpu


def captures(lib, h):
    n = C.c_int64()
    check(lib, lib.fxg_job_graph_captures(h, C.byref(n)), "captures")
    return n.value


def data(spec, seed, **kw):
    """Same delays (so the same frame counts) for every seed."""
    return make_subint(spec, seed=seed, delays_us=[0.37 * k for k in range(spec.nstations)], **kw)


def eager(lib, ctx, job, sd):
    g = GpuJob(lib, ctx, job)
    out = g.process(sd)
    g.close()
    return out


SPECS = {
    "c2": JobSpec(nstations=4, nsubbands=2, npol=2, bw_mhz=32.0, nchan=1024, blockspersend=20),
    "fp16_pcal": JobSpec(nstations=5, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=24,
                         xmac="fp16", pcal_interval_mhz=1.0, base_freq_mhz=8400.01),
    "8bit_chunks": JobSpec(nstations=3, nsubbands=1, npol=2, bw_mhz=16.0, nchan=512, blockspersend=12, nbits=8,
                           max_chunk_blocks=5, zoom=((0, 64, 128),)),
}


@pytest.mark.parametrize("name", sorted(SPECS))
def test_graph_replays_equal_eager(gpu, name):
    lib, ctx = gpu
    spec = SPECS[name]
    job, keep = spec.build()
    g = GpuJob(lib, ctx, job)
    check(lib, lib.fxg_job_set_graph(g.h, 1), "set_graph")
    for seed in (3, 4, 5):
        sd = data(spec, seed, invalid_frames={1: [seed % 3]}, invalid_blocks={0: [seed]})
        got = g.process(sd)
        ref = eager(lib, ctx, job, sd)        # eager work between the replays
        assert np.array_equal(got, ref), (name, seed)
    assert captures(lib, g.h) == 1
    # one more frame of data per station: a new launch shape, a new graph
    sd = data(spec, 6, extra_frames=2)
    got = g.process(sd)
    assert np.array_equal(got, eager(lib, ctx, job, sd))
    assert captures(lib, g.h) == 2
    g.close()


def test_batch_equals_each_job_alone(gpu):
    """Three jobs of different shapes as branches of one graph, two
    subintegrations in a row, device and host result buffers."""
    lib, ctx = gpu
    names = sorted(SPECS)
    jobs = [SPECS[n].build() for n in names]
    hs = [GpuJob(lib, ctx, j) for j, _ in jobs]
    arr = (C.c_void_p * len(hs))(*[h.h.value for h in hs])
    b = C.c_void_p()
    check(lib, lib.fxg_batch_create(ctx, arr, len(hs), C.byref(b)), "batch_create")
    try:
        for rnd, on_dev in ((0, 0), (1, 0), (2, 1)):
            sds = [data(SPECS[n], 10 + rnd + k) for k, n in enumerate(names)]
            subs = (fxcorr.SubintDesc * len(hs))(*[sd.subint_desc() for sd in sds])
            stk = [sd.station_inputs() for sd in sds]
            starr = (C.POINTER(fxcorr.StationInput) * len(hs))(*[C.cast(st, C.POINTER(fxcorr.StationInput))
                                                                 for st, _ in stk])
            outs = [np.zeros(2 * h.info.result_length, np.float32) for h in hs]
            if on_dev:
                dptr = []
                for o in outs:
                    p = C.c_void_p()
                    check(lib, lib.fxg_device_alloc(ctx, o.nbytes, C.byref(p)), "alloc")
                    dptr.append(p)
                res = (C.c_void_p * len(hs))(*[p.value for p in dptr])
            else:
                res = (C.c_void_p * len(hs))(*[o.ctypes.data for o in outs])
            check(lib, lib.fxg_batch_process(b, subs, starr, 0, res, on_dev), "batch_process")
            if on_dev:
                for o, p in zip(outs, dptr):
                    check(lib, lib.fxg_memcpy_d2h(ctx, o.ctypes.data, p, o.nbytes), "d2h")
                    check(lib, lib.fxg_device_free(ctx, p), "free")
            for k, n in enumerate(names):
                ref = eager(lib, ctx, jobs[k][0], sds[k])
                assert np.array_equal(outs[k].view(np.complex64), ref), (n, rnd)
        ncap = C.c_int64()
        check(lib, lib.fxg_batch_graph_captures(b, C.byref(ncap)), "captures")
        assert ncap.value == 2                # host results twice, then device results
    finally:
        lib.fxg_batch_destroy(b)
        for h in hs:
            h.close()


@pytest.mark.parametrize("name", ["c2", "fp16_pcal"])
def test_autotune_keeps_results(gpu, name):
    """Every schedule the tuner tries computes the same result (within 1e-5);
    the one it keeps runs eagerly and as a graph to the same output."""
    lib, ctx = gpu
    spec = SPECS[name]
    job, keep = spec.build()
    g = GpuJob(lib, ctx, job)
    sd = data(spec, 7)
    base = g.process(sd)
    sub = sd.subint_desc()
    st, keep2 = sd.station_inputs()
    res = (fxcorr.TuneResult * 32)()
    n = C.c_int()
    check(lib, lib.fxg_job_autotune(g.h, C.byref(sub), st, 0, 2, res, 32, C.byref(n)), "autotune")
    assert n.value >= 2
    for i in range(n.value):
        assert res[i].accepted and res[i].rel_err <= 1e-5 and res[i].ms > 0, res[i].schedule.as_dict()
    sched = fxcorr.Schedule()
    check(lib, lib.fxg_job_get_schedule(g.h, C.byref(sched)), "get_schedule")
    tuned = g.process(sd)
    rel = np.linalg.norm(tuned - base) / np.linalg.norm(base)
    assert rel <= 1e-5, (sched.as_dict(), rel)
    check(lib, lib.fxg_job_set_graph(g.h, 1), "set_graph")
    assert np.array_equal(g.process(sd), tuned)
    g.close()
