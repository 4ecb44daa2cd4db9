"""Regenerate tests/golden/fixture_hashes.json.

Decodes every Blosc2 frame of the reference's tests/data (written by
c-blosc2 through python-iarray, tests/generator.py:23-38) WITHOUT the oracle:
LZ4 streams go through the system liblz4 (LZ4_decompress_safe, a black box)
and the byte unshuffle is done with numpy, so the hashes are an independent
pin of the oracle and of the GPU decoder.  Run from the repo root:
    python tests/golden/make_fixture_hashes.py [frames-dir ...]
"""
import ctypes
import glob
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def liblz4():
    for name in ("liblz4.so.1", "/opt/conda/lib/liblz4.so.1", "/usr/lib/x86_64-linux-gnu/liblz4.so.1"):
        try:
            L = ctypes.CDLL(name)
            L.LZ4_decompress_safe.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
            return L
        except OSError:
            continue
    raise SystemExit("liblz4 not found")


def decode_frame(data, L):
    hl = struct.unpack(">i", data[11:15])[0]
    nbytes = struct.unpack(">q", data[30:38])[0]
    chunksize = struct.unpack(">i", data[58:62])[0]
    out = bytearray()
    pos = hl
    for _ in range(-(-nbytes // chunksize)):
        ver, _vlz, flags, ts = data[pos:pos + 4]
        cnb, bsz, cb = struct.unpack("<iii", data[pos + 4:pos + 16])
        assert (flags & 0x5) == 0x5 and data[pos + 21] == 1, "expects LZ4 + shuffle in slot 5"
        nblocks = -(-cnb // bsz)
        bstarts = struct.unpack("<%di" % nblocks, data[pos + 32:pos + 32 + 4 * nblocks])
        for j in range(nblocks):
            bsize = min(bsz, cnb - j * bsz)
            nstreams = ts if not (flags & 0x10) and bsize == bsz else 1
            neb = bsize // nstreams
            p = pos + bstarts[j]
            streams = []
            for _s in range(nstreams):
                cs = struct.unpack("<i", data[p:p + 4])[0]
                p += 4
                if cs == 0:
                    streams.append(bytes(neb))
                elif cs == neb:
                    streams.append(data[p:p + cs]); p += cs
                else:
                    buf = ctypes.create_string_buffer(neb)
                    r = L.LZ4_decompress_safe(data[p:p + cs], buf, cs, neb)
                    assert r == neb
                    streams.append(buf.raw); p += cs
            planes = np.frombuffer(b"".join(streams), dtype=np.uint8).reshape(ts, bsize // ts)
            out += planes.T.tobytes()
        pos += cb
    return bytes(out)


def main(dirs):
    L = liblz4()
    res = {}
    for d in dirs:
        for f in sorted(glob.glob(os.path.join(d, "*.iarray"))):
            raw = decode_frame(open(f, "rb").read(), L)
            dt = np.float32 if "float32" in f else np.int32
            arr = np.frombuffer(raw, dtype=dt)
            res[os.path.basename(f)] = {
                "sha256": hashlib.sha256(raw).hexdigest(), "n": int(arr.size),
                "min": float(arr.min()), "max": float(arr.max()), "mean": float(arr.mean())}
    with open(os.path.join(HERE, "fixture_hashes.json"), "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    print(json.dumps({k: v["sha256"][:16] for k, v in res.items()}, indent=1))


if __name__ == "__main__":
    main(sys.argv[1:] or ["/root/reference/tests/data"])
