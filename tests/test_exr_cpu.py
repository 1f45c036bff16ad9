"""The EXR writer of main() (rc_write_exr_image, no GPU): tinyexr's layout
for the reference's SaveEXR call (src/main.cpp:192-207, lib/tinyexr.h
SaveEXR): channels B, G, R as FLOAT, ZIP compression in 16-line blocks for
images of 16 rows or columns and more, none below; decoded by the
spec-following reader in tests/exr_reader.py, the pixels come back bit for bit."""
import os
import sys
import zlib

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from exr_reader import _deinterleave, _unpredict, read_exr, read_rgb  # noqa: E402

from realcausticsengine_amd import engine  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not os.path.exists(engine.LIB_PATH):
        from realcausticsengine_amd import build
        build.build()
    engine.load()


def test_reader_undoes_the_predictor_and_interleave():
    # forward transform written out independently: split even / odd bytes, then byte deltas + 128
    rng = np.random.default_rng(0)
    raw = rng.integers(0, 256, 1001).astype(np.uint8)
    t = np.concatenate([raw[0::2], raw[1::2]])
    d = t.copy()
    d[1:] = (t[1:].astype(np.int64) - t[:-1].astype(np.int64) + 128) & 255
    assert np.array_equal(_deinterleave(_unpredict(d)), raw)


@pytest.mark.parametrize("W,H,comp", [(37, 53, 3), (64, 16, 3), (16, 1, 3), (15, 15, 0), (8, 3, 0)])
def test_exr_roundtrip(tmp_path, W, H, comp):
    rng = np.random.default_rng(W * 100 + H)
    img = rng.random((H, W, 3)).astype(np.float32)
    img[: H // 2, :, 1] = 0.0  # runs of equal values, so ZIP blocks do compress
    img[0, 0] = (np.inf, -0.0, 1e-40)  # special values survive bit for bit
    path = str(tmp_path / "x.exr")
    engine.write_exr_image(path, img)
    e = read_exr(path)
    assert e["compression"] == comp and e["channel_order"] == ["B", "G", "R"]
    assert (e["width"], e["height"]) == (W, H)
    assert e["blocks"] == ((H + 15) // 16 if comp == 3 else H)
    assert set(e["header"]) >= {"channels", "compression", "dataWindow", "displayWindow", "lineOrder",
                               "pixelAspectRatio", "screenWindowCenter", "screenWindowWidth"}
    back = read_rgb(path)
    assert np.array_equal(back.view(np.uint32), img.view(np.uint32))


def test_zip_blocks_are_zlib_streams(tmp_path):
    img = np.zeros((40, 40, 3), dtype=np.float32)
    img[10:20, 5:30] = 0.25
    path = str(tmp_path / "z.exr")
    engine.write_exr_image(path, img)
    e = read_exr(path)
    buf = open(path, "rb").read()
    assert e["compression"] == 3 and len(buf) < 40 * 40 * 12 // 4  # compressed well below the raw size
    for off in e["offsets"]:
        y, size = np.frombuffer(buf[off:off + 8], dtype="<i4")
        assert size < 16 * 40 * 12  # every block of this image compresses
        zlib.decompress(buf[off + 8:off + 8 + size])  # a complete zlib stream
    np.testing.assert_array_equal(read_rgb(path), img)
