"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, keeps the reference struct layouts, and validates input in
init() without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from realcausticsengine_amd import blender, engine, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "real_caustics.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(engine.LIB_PATH):
        from realcausticsengine_amd import build
        build.build()
    return engine.load()


def test_exports_every_declared_symbol(lib):
    declared = _declared_functions()
    assert "init" in declared and "main" in declared
    assert sorted(engine.EXPORTS) == declared
    for name in declared:
        assert hasattr(lib, name), name


def test_struct_layouts():
    assert ctypes.sizeof(blender.PythonLight) == 64
    assert blender.PythonLight.power.offset == 52 and blender.PythonLight.radius.offset == 60
    assert blender.PythonLight.rotation.offset == 40 and blender.PythonLight.color.offset == 4
    assert ctypes.sizeof(blender.PythonMaterial) == 28 and blender.PythonMaterial.color.offset == 16
    assert ctypes.sizeof(blender.DebugStuff) == 24
    assert blender.MVERT_DTYPE.itemsize == 20 and blender.MPOLY_DTYPE.fields["mat_nr"][1] == 8


def test_blender_mesh_block_pointers():
    args = scenes.c0_scene(8, 8, subdiv=1)
    blk = ctypes.string_at(args.meshes_pointers[0], blender.MESH_STRUCT_SIZE)
    mpoly = int.from_bytes(blk[200:208], "little")
    mvert = int.from_bytes(blk[256:264], "little")
    assert mpoly != 0 and mvert != 0
    v0 = np.frombuffer(ctypes.string_at(mvert, 20), dtype=blender.MVERT_DTYPE)[0]
    np.testing.assert_allclose(v0["co"], args.meshes[0].verts[0])


def _init(lib, args):
    lib.init(*args.init_args())
    return lib.rc_last_error().decode()


def test_init_accepts_valid_scene(lib):
    assert _init(lib, scenes.c0_scene(8, 8, subdiv=1)) == ""
    lib.rc_release()


def test_init_rejects_bad_input(lib):
    args = scenes.c0_scene(8, 8, subdiv=1)
    args.camera.width = 0
    assert "resolution" in _init(lib, args)
    args = scenes.c0_scene(8, 8, subdiv=1)
    args.mat_index_arrays[1][0] = 7  # material index beyond the material array
    assert "material index" in _init(lib, args)
    args = scenes.c0_scene(8, 8, subdiv=1)
    args.meshes_pointers[0] = 0
    assert "null Blender mesh" in _init(lib, args)
    args = scenes.c0_scene(8, 8, subdiv=1)
    args.samples = 0
    assert "samples" in _init(lib, args)
    lib.rc_release()


def test_options_roundtrip(lib):
    assert lib.rc_set_option(b"seed", 1234.0) == 0
    assert lib.rc_get_option(b"seed") == 1234.0
    assert lib.rc_set_option(b"no_such_option", 1.0) == -1
    assert lib.rc_set_option_str(b"output_path", b"/tmp/x.exr") == 0
    lib.rc_release()


def test_main_before_init_reports_error(lib):
    lib.rc_release()
    assert lib.main() != 0
    assert "before init" in lib.rc_last_error().decode()


def test_guide_options_and_no_cpu_fallback(lib):
    for key in (b"guide", b"guide_alpha", b"guide_lr", b"guide_batch", b"guide_cells", b"guide_warmup",
                b"guide_train", b"guide_seed"):
        assert lib.rc_set_option(key, 1.0) == 0, key
    assert lib.rc_get_option(b"guide") == 1.0
    assert _init(lib, scenes.c0_scene(8, 8, subdiv=1)) == ""
    # without a GPU every guide call fails and says why; nothing runs on the CPU
    x = np.zeros(6, dtype=np.float32)
    out = np.zeros(engine.GUIDE_BINS, dtype=np.float32)
    rc = lib.rc_guide_forward(x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 1,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert rc == -1
    assert "HIP device" in lib.rc_last_error().decode()
    lib.rc_set_option(b"guide", 0.0)
    lib.rc_release()
