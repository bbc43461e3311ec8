"""datastream.py -- CPU restatement of the station data window computation.

TEST INFRASTRUCTURE ONLY: tests/ use it to check the host library's
fxds_control; nothing in the product imports it.

Restates, for one buffer holding the whole stream (the reference's ring of
segments reduced to a single segment, whose "next segment" is itself):
  DataStream::calculateControlParams   mpifxcorr/src/datastream.cpp:376-630
  DataStream::updateConfig (bytesbetweenintegerns, sampletimens)  :745-771
  Mk5DataStream::calculateControlParams (frame boundary)  mk5.cpp:146-268
  Mk5DataStream::updateConfig (nsinc)                     mk5.cpp:270-290
C integer semantics (truncating division and remainder toward zero) are
spelled out where Python's differ.
"""
from __future__ import annotations

import math

INVALID_SUBINT = -99999999          # mode.h:232
FLAGS_PER_INT = 30                  # mpifxcorr.h:39
TINY = 0.000000001                  # mode.cpp:35


def cdiv(a: int, b: int) -> int:
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def cmod(a: int, b: int) -> int:
    return a - cdiv(a, b) * b


def ctrunc(x: float) -> int:
    return int(math.trunc(x))


class Stream:
    def __init__(self, framebytes, payloadbytes, framespersecond, framegranularity, bpsnum, bpsden, numchannels,
                 blockspersend, bw_mhz, databytes):
        self.framebytes, self.payloadbytes = framebytes, payloadbytes
        self.fps, self.gran = framespersecond, max(1, framegranularity)
        self.num, self.den = bpsnum, bpsden
        self.numchannels, self.bps = numchannels, blockspersend
        self.sampletimens = 500.0 / bw_mhz
        self.databytes = databytes
        acc, self.bbin = 0.0, 0
        while True:
            acc += bpsden * self.sampletimens
            self.bbin += bpsnum
            if abs(acc - ctrunc(acc + 0.5)) < TINY:
                break

    def set_buffer(self, nbytes, scan, sec, ns):
        self.bytes, self.scan, self.sec, self.ns = nbytes, scan, sec, ns
        self.nsinc = ctrunc((nbytes // self.framebytes) * (1000000000.0 / self.fps) + 0.5)

    def control(self, scan, offsetsec, offsetns, delayus1, delayus2, model_ok=True):
        """(sec, ns, flags list, bufferindex); sec = INVALID_SUBINT if torched."""
        nwords = cdiv(self.bps + FLAGS_PER_INT - 1, FLAGS_PER_INT)
        flags = [0] * nwords
        bad = (INVALID_SUBINT, 0, [0] * nwords, 0)
        if not model_ok:
            return bad
        readbytes = bufferbytes = validbytes = self.bytes
        firstoffsetns = offsetns - ctrunc(delayus1 * 1000)
        dataspanns = ctrunc(self.numchannels * self.bps * 2 * self.sampletimens + 0.5)
        lastoffsetns = offsetns + dataspanns + ctrunc(-1000 * delayus2 + 0.5)
        for _ in range(2):
            if lastoffsetns < 0:
                offsetsec -= 1
                firstoffsetns += 1000000000
                lastoffsetns += 1000000000
        if scan < self.scan or readbytes <= 0:
            return bad
        if scan == self.scan and offsetsec < self.sec - 3:
            return bad
        nsdiff = (offsetsec - self.sec) * 1000000000 + firstoffsetns - self.ns
        validns = cdiv(validbytes * self.nsinc, readbytes)
        if scan > self.scan or nsdiff >= validns:
            return bad
        if nsdiff < -self.nsinc:
            return bad
        blockbytes = cdiv(self.numchannels * 2 * self.num, self.den)
        samples = ctrunc(((offsetsec - self.sec) * 1000000000 + (firstoffsetns - self.ns)) / self.sampletimens + 0.5)
        bufferindex = cdiv(samples * self.num, self.den)
        if cmod(bufferindex, 2) != 0:
            bufferindex -= 1
        if bufferindex == bufferbytes:
            return bad
        count = 0
        while bufferindex < 0 and count < self.bps:
            bufferindex += blockbytes
            count += 1
        if bufferindex < 0:
            return bad
        if count > 0:
            def subtract(c):
                return cdiv(ctrunc(c * 1000.0 * (delayus2 - delayus1) / (self.sampletimens * self.bps) + 0.5) * self.num,
                            self.den)
            if bufferindex - subtract(count) < 0:
                count += 1
                bufferindex += blockbytes
                if count == self.bps:
                    return bad
                bufferindex -= subtract(count)
            bufferindex -= cmod(bufferindex, 2)
        segoff = bufferindex
        bufferindex -= cmod(segoff, self.bbin)
        segoff -= cmod(segoff, self.bbin)
        if validbytes - segoff >= self.databytes:
            for i in range(count, self.bps):
                flags[i // FLAGS_PER_INT] |= 1 << (i % FLAGS_PER_INT)
        else:
            dbi = ((((offsetsec - self.sec) * 1.0e9 + (firstoffsetns - self.ns)) / self.sampletimens) * self.num) / self.den
            if dbi < -blockbytes * self.bps or dbi > bufferbytes:
                return bad
            for i in range(count, self.bps):
                if bufferindex + i * blockbytes < validbytes:
                    flags[i // FLAGS_PER_INT] |= 1 << (i % FLAGS_PER_INT)
            if not any(flags[i] > 0 for i in range(self.bps // FLAGS_PER_INT)):
                return bad
        framesin = cdiv(bufferindex, self.payloadbytes)
        if cmod(framesin, self.gran) != 0:
            framesin -= cmod(framesin, self.gran)
            if framesin < 0:
                framesin += self.gran
        segoffns = self.ns + ctrunc((1000000000.0 * framesin) / self.fps + 0.5)
        sec = self.sec + cdiv(segoffns, 1000000000)
        ns = cmod(segoffns, 1000000000)
        bufferindex = framesin * self.framebytes
        if bufferindex >= bufferbytes:
            return bad
        return sec, ns, flags, bufferindex
