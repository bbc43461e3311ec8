"""Asynchronous submission (fxg_subint_submit / query / wait) and the
request loop (include/fxcorr_serve.h) on the GPU.

  - a submitted subintegration gives the same results as fxg_subint_process;
  - host inputs go through two raw buffer sets on a copy stream: submissions
    of alternating data, eagerly and as graph launches, each equal the same
    data processed alone with its raw bytes already on the device;
  - several configurations of different sizes, with requests interleaved so
    that work of different configurations and slots overlaps (with no bound
    on the requests on the device at once, and with one or two started in
    submission order): every result equals the result of the same
    subintegration processed alone, bit for bit;
  - admission by HBM footprint: a configuration that does not fit the budget
    is refused, and a configuration's slots never exceed its limit or the
    budget.
"""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import FXG_ENOMEM, JobSpec, PlanInfo, ServeStats, SubintDesc, check
from fxsynth import make_subint
from fxtestlib import GpuJob


class DeviceSubint:
    """A subintegration's raw data in device memory and its StationInput array."""

    def __init__(self, lib, ctx, sd):
        self.lib, self.ctx = lib, ctx
        self.ptrs = []
        for st in sd.stations:
            p = C.c_void_p()
            check(lib, lib.fxg_device_alloc(ctx, st.raw.size, C.byref(p)), "alloc")
            check(lib, lib.fxg_memcpy_h2d(ctx, p, st.raw.ctypes.data, st.raw.size), "h2d")
            self.ptrs.append(p.value)
        self.st, self.keep = sd.station_inputs(self.ptrs)
        self.sub = sd.subint_desc()

    def free(self):
        for p in self.ptrs:
            self.lib.fxg_device_free(self.ctx, C.c_void_p(p))


def device_results(lib, ctx, n):
    p = C.c_void_p()
    check(lib, lib.fxg_device_alloc(ctx, 8 * n, C.byref(p)), "alloc results")
    return p


def fetch(lib, ctx, p, n):
    out = np.zeros(n, dtype=np.complex64)
    check(lib, lib.fxg_memcpy_d2h(ctx, out.ctypes.data, p, 8 * n), "d2h")
    return out


SPECS = [JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=128, blockspersend=16, xmac="fp16"),
         JobSpec(nstations=6, nsubbands=2, npol=2, bw_mhz=32.0, nchan=512, blockspersend=8, xmac="fp16"),
         JobSpec(nstations=10, nsubbands=1, npol=2, bw_mhz=32.0, nchan=1024, blockspersend=6, xmac="fp32")]


@pytest.mark.gpu
def test_submit_wait_equals_process(gpu):
    lib, ctx = gpu
    spec = SPECS[0]
    job, keep = spec.build()
    sd = make_subint(spec, seed=41, invalid_frames={1: [2]})
    g = GpuJob(lib, ctx, job)
    ref = g.process(sd)
    st, k2 = sd.station_inputs()
    sub = sd.subint_desc()
    outs, tickets = [], []
    for _ in range(3):
        out = np.zeros(ref.size, dtype=np.complex64)
        t = C.c_int64()
        check(lib, lib.fxg_subint_submit(g.h, C.byref(sub), st, 0, out.ctypes.data, 0, C.byref(t)), "submit")
        outs.append(out)
        tickets.append(t.value)
    assert tickets == sorted(tickets)
    check(lib, lib.fxg_subint_wait(g.h, tickets[-1]), "wait")
    done = C.c_int()
    for t in tickets:
        check(lib, lib.fxg_subint_query(g.h, t, C.byref(done)), "query")
        assert done.value == 1
    for out in outs:
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    assert lib.fxg_subint_wait(g.h, tickets[-1] + 5) == fxcorr.FXG_EINVAL
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [0, 1])
def test_host_inputs_pipelined(gpu, graph):
    """Host raw data copied on the job's copy stream into one of two buffer
    sets while the previous submission's kernels run: six submissions of
    three different subintegrations in turn (a set is refilled while the
    other is read), bit for bit equal to device-resident inputs."""
    lib, ctx = gpu
    spec = SPECS[1]
    job, keep = spec.build()
    sds = [make_subint(spec, seed=60 + k, invalid_frames={k: [1]}) for k in range(3)]
    g = GpuJob(lib, ctx, job)
    n = g.info.result_length
    refs = []
    for sd in sds:
        dev = DeviceSubint(lib, ctx, sd)
        res = device_results(lib, ctx, n)
        check(lib, lib.fxg_subint_process(g.h, C.byref(dev.sub), dev.st, 1, res, 1), "process")
        refs.append(fetch(lib, ctx, res, n))
        lib.fxg_device_free(ctx, res)
        dev.free()
    check(lib, lib.fxg_job_set_graph(g.h, graph), "graph")
    ins = [sd.station_inputs() for sd in sds]
    subs = [sd.subint_desc() for sd in sds]
    outs, tickets = [], []
    for r in range(6):
        k = r % 3
        out = np.zeros(n, dtype=np.complex64)
        t = C.c_int64()
        check(lib, lib.fxg_subint_submit(g.h, C.byref(subs[k]), ins[k][0], 0, out.ctypes.data, 0, C.byref(t)),
              "submit")
        outs.append((k, out))
        tickets.append(t.value)
    check(lib, lib.fxg_subint_wait(g.h, tickets[-1]), "wait")
    for k, out in outs:
        assert np.array_equal(out.view(np.uint32), refs[k].view(np.uint32)), k
    assert not np.array_equal(refs[0], refs[1])
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("inflight", [0, 1, 2])
def test_results_alone_equal_results_in_the_mix(gpu, inflight):
    lib, ctx = gpu
    host = fxcorr.load_host_lib()
    jobs, alone, subs, lens = [], [], [], []
    for i, spec in enumerate(SPECS):
        job, keep = spec.build()
        jobs.append((job, keep))
        per = []
        for k in range(2):                      # two different subintegrations per configuration
            sd = make_subint(spec, seed=100 + 10 * i + k, invalid_frames={0: [1 + k]})
            g = GpuJob(lib, ctx, job)
            per.append((sd, g.process(sd)))
            lens.append(g.info.result_length)
            g.close()
        alone.append(per)
        subs.append([DeviceSubint(lib, ctx, sd) for sd, _ in per])
    srv = C.c_void_p()
    check(host, host.fxs_create(0, 0, C.byref(srv)), "fxs_create")
    try:
        check(host, host.fxs_set_max_inflight(srv, inflight), "max_inflight")
        cids = []
        for job, keep in jobs:
            cid = C.c_int()
            check(host, host.fxs_add_config(srv, C.byref(job), 2, C.byref(cid)), "add_config")
            cids.append(cid.value)
        pending = []
        for rnd in range(3):                    # interleave the configurations, several in flight
            for i in range(len(SPECS)):
                for k in range(2):
                    n = alone[i][k][1].size
                    res = device_results(lib, ctx, n)
                    req = C.c_int64()
                    ds = subs[i][k]
                    check(host, host.fxs_submit(srv, cids[i], C.byref(ds.sub), ds.st, res, C.byref(req)), "submit")
                    pending.append((req.value, i, k, res, n))
        lat = C.c_int64()
        for req, i, k, res, n in pending:
            check(host, host.fxs_wait(srv, req, C.byref(lat)), "wait")
            assert lat.value > 0
            got = fetch(lib, ctx, res, n)
            assert np.array_equal(got.view(np.uint32), alone[i][k][1].view(np.uint32)), (i, k)
            lib.fxg_device_free(ctx, res)
        stats = ServeStats()
        check(host, host.fxs_stats(srv, C.byref(stats)), "stats")
        assert stats.completed == stats.submitted == len(pending)
        assert stats.queued == 0
        assert 3 <= stats.nslots <= 6 and stats.used_bytes <= stats.budget_bytes
    finally:
        host.fxs_destroy(srv)
        for per in subs:
            for ds in per:
                ds.free()


@pytest.mark.gpu
def test_admission_by_hbm_footprint(gpu):
    lib, ctx = gpu
    host = fxcorr.load_host_lib()
    small, keep1 = SPECS[0].build()
    big, keep2 = SPECS[2].build()
    info_s, info_b = PlanInfo(), PlanInfo()
    check(lib, lib.fxg_plan(C.byref(small), C.byref(info_s)), "plan")
    check(lib, lib.fxg_plan(C.byref(big), C.byref(info_b)), "plan")
    assert info_b.device_bytes > info_s.device_bytes
    budget = info_s.device_bytes + info_s.device_bytes // 2   # one small slot fits, two do not
    srv = C.c_void_p()
    check(host, host.fxs_create(0, budget, C.byref(srv)), "fxs_create")
    try:
        cid = C.c_int()
        assert host.fxs_add_config(srv, C.byref(big), 1, C.byref(cid)) == FXG_ENOMEM
        check(host, host.fxs_add_config(srv, C.byref(small), 4, C.byref(cid)), "add small")
        sd = make_subint(SPECS[0], seed=7)
        ds = DeviceSubint(lib, ctx, sd)
        n = info_s.result_length
        reqs = []
        for _ in range(4):
            res = device_results(lib, ctx, n)
            req = C.c_int64()
            check(host, host.fxs_submit(srv, cid.value, C.byref(ds.sub), ds.st, res, C.byref(req)), "submit")
            reqs.append((req.value, res))
        outs = []
        for req, res in reqs:
            check(host, host.fxs_wait(srv, req, None), "wait")
            outs.append(fetch(lib, ctx, res, n))
            lib.fxg_device_free(ctx, res)
        assert all(np.array_equal(o.view(np.uint32), outs[0].view(np.uint32)) for o in outs)
        stats = ServeStats()
        check(host, host.fxs_stats(srv, C.byref(stats)), "stats")
        assert stats.nslots == 1 and stats.used_bytes <= budget
        ds.free()
    finally:
        host.fxs_destroy(srv)
