"""Multi-thread VDIF streams for the multiplexer tests, and runners for the
oracle's and the library's vdifmux."""
from __future__ import annotations

import ctypes as C

import numpy as np

import fxcorr

FILL = np.array([0x11223344], dtype=np.uint32).view(np.uint8)


def header(sec, fn, thread, framebytes, nchan, nbits, invalid=False, epoch=20, station=7):
    w = np.zeros(8, dtype=np.uint32)
    w[0] = (sec & 0x3FFFFFFF) | ((1 << 31) if invalid else 0)
    w[1] = (fn & 0xFFFFFF) | (epoch << 24)
    w[2] = ((framebytes // 8) & 0xFFFFFF) | ((int(np.log2(nchan)) & 0x1F) << 24)
    w[3] = (station & 0xFFFF) | ((thread & 0x3FF) << 16) | (((nbits - 1) & 0x1F) << 26)
    return w.view(np.uint8)


def thread_stream(ids, nframes, framebytes=1032, fps=1000, nbits=2, nchan=1, seed=1, sec0=100, first=0,
                  shuffle=0, drop=(), dup=(), invalid=(), fill_after=(), junk_after=(), wrong_thread_after=()):
    """Frames of len(ids) threads, frame numbers first..first+nframes-1, each
    set of thread frames in time order; shuffle > 0 permutes frames within
    windows of that many.  drop/dup/invalid: (frame, thread index) pairs;
    fill_after / junk_after / wrong_thread_after: frame indices after which a
    fill-pattern frame, 12 junk bytes or a frame of an unwanted thread go."""
    rng = np.random.default_rng(seed)
    out, payloads = [], {}
    for f in range(first, first + nframes):
        for k, tid in enumerate(ids):
            p = rng.integers(0, 256, framebytes - 32, dtype=np.uint8)
            payloads[(f, k)] = p
            if (f, k) in drop:
                continue
            fr = np.concatenate([header(sec0 + f // fps, f % fps, tid, framebytes, nchan, nbits,
                                        invalid=(f, k) in invalid), p])
            out.append(fr)
            if (f, k) in dup:
                out.append(fr.copy())
        if f in fill_after:
            out.append(np.tile(FILL, framebytes // 4))
        if f in junk_after:
            out.append(rng.integers(0, 256, 12, dtype=np.uint8))
        if f in wrong_thread_after:
            out.append(np.concatenate([header(sec0 + f // fps, f % fps, 999, framebytes, nchan, nbits),
                                       rng.integers(0, 256, framebytes - 32, dtype=np.uint8)]))
    if shuffle > 1:
        for a in range(0, len(out), shuffle):
            seg = out[a:a + shuffle]
            order = rng.permutation(len(seg))
            out[a:a + shuffle] = [seg[i] for i in order]
    return np.concatenate(out), payloads


def orc_mux(orc, src, ids, dest_frames, framebytes=1032, fps=1000, nbits=2, nchan=1, nsort=8, ngap=8, flags=0,
            start=-1):
    nt = len(ids)
    units = 1
    while units < nt:
        units *= 2
    outframe = (framebytes - 32) * units + 32
    dest = np.zeros(outframe * dest_frames, dtype=np.uint8)
    st = (C.c_int64 * 18)()
    idarr = (C.c_int * nt)(*ids)
    src = np.ascontiguousarray(src)
    r = orc.orc_vdifmux(dest.ctypes.data, dest.size, src.ctypes.data, src.size, framebytes, fps, nbits, nt, idarr,
                        nsort, ngap, flags, nchan, start, st)
    return r, dest, list(st)


def mux_desc(ids, framebytes=1032, fps=1000, nbits=2, nchan=1, nsort=8, ngap=8, flags=0):
    d = fxcorr.VdifMuxDesc()
    d.input_frame_bytes, d.frames_per_second, d.bits_per_sample = framebytes, fps, nbits
    d.nthreads = len(ids)
    for i, t in enumerate(ids):
        d.thread_ids[i] = t
    d.channels_per_thread, d.nsort, d.ngap, d.flags = nchan, nsort, ngap, flags
    return d
