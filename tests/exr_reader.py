"""A small OpenEXR reader for the tests (single-part scanline files with
FLOAT, HALF or UINT channels, NO_COMPRESSION or ZIPS/ZIP compression), written
from the OpenEXR file layout: magic 20000630, version 2, attributes (name,
type, size, value) up to an empty name, one uint64 offset per block, and per
block the first scanline, the data size and the data.  ZIP data is a zlib
stream of the block's bytes after OpenEXR's predictor (byte deltas + 128)
and interleaving (even bytes first, then odd bytes); a block whose size equals
its raw size is stored raw."""
from __future__ import annotations

import struct
import zlib

import numpy as np

_PIXEL = {0: ("<u4", 4), 1: ("<f2", 2), 2: ("<f4", 4)}


def _unpredict(t: np.ndarray) -> np.ndarray:
    out = np.empty_like(t)
    acc = int(t[0])
    out[0] = acc
    d = t[1:].astype(np.int64) - 128
    out[1:] = (acc + np.cumsum(d)) & 255
    return out


def _deinterleave(t: np.ndarray) -> np.ndarray:
    n = len(t)
    h = (n + 1) // 2
    out = np.empty_like(t)
    out[0::2] = t[:h]
    out[1::2] = t[h:]
    return out


def read_exr(path: str) -> dict:
    """Returns {"channels": {name: (H, W) array}, "compression": int, "width", "height", "header": {...}}."""
    buf = open(path, "rb").read()
    magic, version = struct.unpack_from("<II", buf, 0)
    if magic != 20000630:
        raise ValueError("not an OpenEXR file")
    if version & 0xFF != 2 or version & 0x200:
        raise ValueError("only single-part scanline files")
    pos = 8
    header = {}
    while True:
        end = buf.index(b"\0", pos)
        name = buf[pos:end].decode()
        pos = end + 1
        if not name:
            break
        end = buf.index(b"\0", pos)
        typ = buf[pos:end].decode()
        pos = end + 1
        (size,) = struct.unpack_from("<I", buf, pos)
        pos += 4
        header[name] = (typ, buf[pos:pos + size])
        pos += size
    chans = []
    data = header["channels"][1]
    p = 0
    while data[p] != 0:
        end = data.index(b"\0", p)
        cname = data[p:end].decode()
        ptype, _, xs, ys = struct.unpack_from("<iiii", data, end + 1)
        chans.append((cname, ptype))
        p = end + 1 + 16
    comp = header["compression"][1][0]
    x0, y0, x1, y1 = struct.unpack("<iiii", header["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    lines = {0: 1, 2: 1, 3: 16}[comp]
    blocks = (H + lines - 1) // lines
    offsets = struct.unpack_from(f"<{blocks}Q", buf, pos)
    planes = {c: np.zeros((H, W), dtype=_PIXEL[t][0]) for c, t in chans}
    line_bytes = sum(_PIXEL[t][1] for _, t in chans) * W
    for off in offsets:
        y, size = struct.unpack_from("<iI", buf, off)
        n = min(lines, y1 + 1 - y)
        raw_size = line_bytes * n
        blob = buf[off + 8:off + 8 + size]
        if comp == 0 or size == raw_size:
            raw = np.frombuffer(blob, dtype=np.uint8)
        else:
            t = np.frombuffer(zlib.decompress(blob), dtype=np.uint8)
            if len(t) != raw_size:
                raise ValueError("ZIP block of the wrong size")
            raw = _deinterleave(_unpredict(t))
        q = 0
        for k in range(n):
            for c, t in chans:
                dt, sz = _PIXEL[t]
                planes[c][y - y0 + k] = np.frombuffer(raw[q:q + sz * W].tobytes(), dtype=dt)
                q += sz * W
    return {"channels": planes, "compression": comp, "width": W, "height": H, "header": header,
            "channel_order": [c for c, _ in chans], "blocks": blocks, "offsets": offsets}


def read_rgb(path: str) -> np.ndarray:
    """(H, W, 3) float32 image of an R, G, B file."""
    e = read_exr(path)
    return np.stack([e["channels"][c].astype(np.float32) for c in ("R", "G", "B")], axis=-1)
