"""fxg_unpack_noalloc (the unpack entry point that allocates nothing: its
frame and station tables live in the library's static device scratch)
against fxg_unpack: decoded samples bit for bit and the same good-sample
weight, on two calls in a row of which the second has the smaller input, for
VDIF 2- and 8-bit and Mark5B streams with invalid and fill-pattern frames;
a stream of more frames than the scratch holds is refused before any device
work."""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import FXG_EINVAL, FXG_FORMAT_MARK5B, FXG_FORMAT_VDIF, check
from test_gpu_parity import ds_desc, random_frames

pytestmark = pytest.mark.gpu

SCRATCH_FRAMES = ((4 << 20) - 512) // 8

CASES = [(FXG_FORMAT_VDIF, 2, 2, 8032, 32), (FXG_FORMAT_VDIF, 4, 8, 8032, 32), (FXG_FORMAT_MARK5B, 8, 2, 10016, 16)]


def unpack(lib, ctx, fn, d, dptr, nbytes, off, nsamp, nchan):
    out = C.c_void_p()
    check(lib, lib.fxg_device_alloc(ctx, 4 * nchan * nsamp, C.byref(out)), "alloc")
    w = C.c_float()
    try:
        check(lib, getattr(lib, fn)(ctx, C.byref(d), dptr, nbytes, off, nsamp, out, C.byref(w)), fn)
        host = np.zeros(nchan * nsamp, dtype=np.float32)
        check(lib, lib.fxg_memcpy_d2h(ctx, host.ctypes.data, out, host.nbytes), "d2h")
    finally:
        lib.fxg_device_free(ctx, out)
    return host, w.value


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"fmt{c[0]}-{c[1]}ch-{c[2]}bit")
def test_noalloc_equals_unpack_twice(gpu, case):
    lib, ctx = gpu
    fmt, nchan, nbits, fb, hb = case
    rng = np.random.default_rng(900 + nchan + nbits)
    big = random_frames(rng, fmt, nchan, nbits, 40, fb, hb,
                        [(3, "invalid"), (9, "fillstart"), (17, "fillend"), (25, "filltail")])
    small = random_frames(rng, fmt, nchan, nbits, 12, fb, hb, [(2, "invalid"), (7, "fillhead")])
    d = ds_desc(fmt, nchan, nbits, fb, hb)
    fs = (fb - hb) * 8 // (nchan * nbits)
    # each window covers an invalid frame (big: 3 and 9; small: 2), so the weight is below 1
    for data, off, nsamp in ((big, 3 * fs - 77, 8 * fs + 5), (small, 2 * fs - 100, 4096)):
        print(f"case {case}: {data.size} bytes, offset {off}, {nsamp} samples", flush=True)
        dptr = C.c_void_p()
        check(lib, lib.fxg_device_alloc(ctx, data.size, C.byref(dptr)), "alloc")
        try:
            check(lib, lib.fxg_memcpy_h2d(ctx, dptr, data.ctypes.data, data.size), "h2d")
            ref, wref = unpack(lib, ctx, "fxg_unpack", d, dptr, data.size, off, nsamp, nchan)
            got, w = unpack(lib, ctx, "fxg_unpack_noalloc", d, dptr, data.size, off, nsamp, nchan)
        finally:
            lib.fxg_device_free(ctx, dptr)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        assert w == wref and 0.0 < w < 1.0


def test_noalloc_refuses_more_frames_than_the_scratch(gpu):
    lib, ctx = gpu
    fmt, nchan, nbits, fb, hb = CASES[0]
    d = ds_desc(fmt, nchan, nbits, fb, hb)
    dptr = C.c_void_p()
    check(lib, lib.fxg_device_alloc(ctx, fb, C.byref(dptr)), "alloc")
    out = C.c_void_p()
    check(lib, lib.fxg_device_alloc(ctx, 4 * nchan * 64, C.byref(out)), "alloc")
    w = C.c_float()
    try:
        # the length is checked against the scratch before any device work
        rc = lib.fxg_unpack_noalloc(ctx, C.byref(d), dptr, (SCRATCH_FRAMES + 1) * fb, 0, 64, out, C.byref(w))
        assert rc == FXG_EINVAL
        assert b"scratch" in lib.fxg_last_error()
    finally:
        lib.fxg_device_free(ctx, out)
        lib.fxg_device_free(ctx, dptr)
