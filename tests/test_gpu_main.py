"""The plugin's call sequence end to end: init(...) then main() with
debug_test = 1 renders `samples` iterations, averages them and writes the
EXR (src/main.cpp:89-223); the plugin reads that file.  The decoded file
must equal the same render through rc_render_iteration / rc_get_image, and
match the CPU oracle's render within the parity bar (2e-4 relative + 1e-7
absolute per pixel, <= 0.2 % of pixels outside: the merge sums in another
order).  main() also releases the scene, as the reference does."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from exr_reader import read_exr, read_rgb  # noqa: E402

from realcausticsengine_amd import scenes  # noqa: E402
from realcausticsengine_amd.engine import Engine, EngineError  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h", [(48, 40), (12, 10)])
def test_main_writes_the_rendered_exr(oracle, tmp_path, w, h):
    args = scenes.c0_scene(w, h, subdiv=3, samples=2)
    opts = dict(photons=16384, seed=11)
    # the same render through the per-iteration calls
    eng = Engine(args, write_exr=0, **opts)
    eng.render_iteration()
    eng.render_iteration()
    direct = eng.image(scaled=True)
    eng.release()
    out = str(tmp_path / "caustics.exr")
    eng = Engine(args, output_path=out, **opts)
    assert eng.main() == 0
    e = read_exr(out)
    assert e["compression"] == (3 if max(w, h) >= 16 else 0)
    img = read_rgb(out)
    assert img.shape == (h, w, 3)
    same = np.all(img.view(np.uint32) == direct.view(np.uint32), axis=2)
    assert same.mean() >= 0.999, same.mean()  # film atomics may reorder a shared pixel's adds
    ref = oracle.OracleScene(args, **opts).render(2)
    bad = np.abs(img - ref) > 2e-4 * np.abs(ref) + 1e-7
    print(f"{w}x{h}: sum {img.sum():.6g} vs oracle {ref.sum():.6g}, outside {bad.mean():.4f}", flush=True)
    assert ref.sum() > 0 and bad.mean() <= 0.002
    # main() released the scene: the next render needs init again
    with pytest.raises(EngineError):
        eng.render_iteration()
