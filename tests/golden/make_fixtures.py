"""Writes the committed fixtures under tests/golden/ from data files held in
the reference tree (run in a container where the reference is mounted):

  ma008_1_tables.json  frequency / datastream / baseline tables of the
                       reference job sites/MPIfR/oneoff/difx2difx_testset/
                       ma008_1.input (a real 3-station INTERLACEDVDIF job)
  vdifio_1.vdif        the VDIF sample file libraries/vdifio/python/examples/
                       vdifio_1.vdif (10 frames, 4 channels, 2 bits)
  vdifio_1_decoded.npz its samples decoded by an independent bit-level
                       reading of the VDIF layout (sample k = t*nchan + c at
                       bits [2k, 2k+2) of the payload, code order
                       -HiMag,-1,+1,+HiMag), the expected output of unpack
"""
import json
import os
import shutil
import struct

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_input(path):
    lines = [l.rstrip("\n") for l in open(path)]
    kv = []
    for l in lines:
        if ":" in l and not l.startswith("#"):
            k, v = l.split(":", 1)
            kv.append((k.strip(), v.strip()))
    it = iter(kv)
    cur = [None]

    def take(prefix):
        for k, v in it:
            if k.startswith(prefix):
                return v
        raise KeyError(prefix)

    freqs = []
    n = int(take("FREQ ENTRIES"))
    for i in range(n):
        f = float(take(f"FREQ (MHZ) {i}"))
        bw = float(take(f"BW (MHZ) {i}"))
        sb = take(f"SIDEBAND {i}")
        nch = int(take(f"NUM CHANNELS {i}"))
        avg = int(take(f"CHANS TO AVG {i}"))
        freqs.append({"freq_mhz": f, "bw_mhz": bw, "lsb": sb == "L", "nchan": nch, "chans_to_avg": avg})
    dss = []
    n = int(take("DATASTREAM ENTRIES"))
    for d in range(n):
        fmt = take("DATA FORMAT")
        nbits = int(take("QUANTISATION BITS"))
        framebytes = int(take("DATA FRAME SIZE"))
        nrf = int(take("NUM RECORDED FREQS"))
        recf = []
        for i in range(nrf):
            fi = int(take(f"REC FREQ INDEX {i}"))
            clk = float(take(f"CLK OFFSET {i}"))
            fo = float(take(f"FREQ OFFSET {i}"))
            npol = int(take(f"NUM REC POLS {i}"))
            recf.append({"index": fi, "clock_us": clk, "looffset_hz": fo, "npol": npol})
        nb = sum(r["npol"] for r in recf)
        recb = []
        for b in range(nb):
            pol = take(f"REC BAND {b} POL")
            idx = int(take(f"REC BAND {b} INDEX"))
            recb.append({"pol": pol, "localfreq": idx})
        nzf = int(take("NUM ZOOM FREQS"))
        zf = []
        for i in range(nzf):
            fi = int(take(f"ZOOM FREQ INDEX {i}"))
            npol = int(take(f"NUM ZOOM POLS {i}"))
            zf.append({"index": fi, "npol": npol})
        nzb = sum(z["npol"] for z in zf)
        zb = []
        for b in range(nzb):
            pol = take(f"ZOOM BAND {b} POL")
            idx = int(take(f"ZOOM BAND {b} INDEX"))
            zb.append({"pol": pol, "localfreq": idx})
        dss.append({"format": fmt, "nbits": nbits, "framebytes": framebytes, "recfreqs": recf,
                    "recbands": recb, "zoomfreqs": zf, "zoombands": zb})
    bls = []
    n = int(take("BASELINE ENTRIES"))
    for b in range(n):
        a = int(take(f"D/STREAM A INDEX {b}"))
        c = int(take(f"D/STREAM B INDEX {b}"))
        nf = int(take(f"NUM FREQS {b}"))
        fl = []
        for f in range(nf):
            npp = int(take(f"POL PRODUCTS {b}/{f}"))
            prods = []
            for p in range(npp):
                b1 = int(take(f"D/STREAM A BAND {p}"))
                b2 = int(take(f"D/STREAM B BAND {p}"))
                prods.append([b1, b2])
            fl.append(prods)
        bls.append({"ds1": a, "ds2": c, "freqs": fl})
    return {"freqs": freqs, "datastreams": dss, "baselines": bls}


def decode_vdif_bits(raw):
    """Independent bit-level decode of a VDIF file with 32-byte headers."""
    out = []
    pos = 0
    levels = np.array([-3.3359, -1.0, 1.0, 3.3359], dtype=np.float32)
    while pos + 32 <= len(raw):
        w = struct.unpack("<8I", raw[pos:pos + 32])
        flen = (w[2] & 0xFFFFFF) * 8
        nchan = 1 << ((w[2] >> 24) & 0x1F)
        nbits = ((w[3] >> 26) & 0x1F) + 1
        assert nbits == 2
        payload = np.frombuffer(raw[pos + 32: pos + flen], dtype=np.uint8)
        bits = np.unpackbits(payload, bitorder="little").reshape(-1, 2)
        codes = bits[:, 0] + 2 * bits[:, 1]
        out.append(levels[codes].reshape(-1, nchan))
        pos += flen
    return np.concatenate(out, axis=0), nchan


def main():
    t = parse_input(os.path.join(REF, "sites/MPIfR/oneoff/difx2difx_testset/ma008_1.input"))
    json.dump(t, open(os.path.join(HERE, "ma008_1_tables.json"), "w"), indent=1)
    src = os.path.join(REF, "libraries/vdifio/python/examples/vdifio_1.vdif")
    shutil.copyfile(src, os.path.join(HERE, "vdifio_1.vdif"))
    raw = open(src, "rb").read()
    dec, nchan = decode_vdif_bits(raw)
    np.savez_compressed(os.path.join(HERE, "vdifio_1_decoded.npz"), samples=dec.T.copy(), nchan=nchan)
    print("fixtures written:", dec.shape, nchan)


if __name__ == "__main__":
    main()
