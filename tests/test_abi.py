"""The C ABI library loads and exports every symbol include/vgi.h declares
(no GPU calls)."""
import ctypes
import os
import re

from conftest import ROOT, load_package


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vgi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vgi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    vgi = load_package()
    lib = vgi.library()
    names = declared_symbols()
    assert len(names) > 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # The Python layer binds exactly the declared surface.
    assert sorted(vgi.exported_symbols()) == names


def test_struct_layouts_match_the_reference():
    """gpu.rs:1367-1371 / voxel_trace.wgsl:5-20: 32-byte tasks, 96-byte
    results; GpuLight is 48 bytes (gpu.rs:136-141)."""
    vgi = load_package()
    assert vgi.FACE_TASK.itemsize == 32
    assert ctypes.sizeof(vgi.Light) == 48


def test_errors_surface_as_messages():
    vgi = load_package()
    lib = vgi.library()
    assert lib.vgi_set_transport(None, None) < 0
    assert b"null" in lib.vgi_last_error()


def test_every_entry_point_rejects_a_null_handle():
    """Every entry point taking a tracer or batch handle returns a negative
    code and leaves a message when the handle is NULL (no GPU call is made);
    the float and count queries answer 0."""
    vgi = load_package()
    lib = vgi.library()
    checked = 0
    for name in vgi.exported_symbols():
        if name.startswith("vgi_scene_") or name in ("vgi_last_error", "vgi_destroy", "vgi_batch_destroy"):
            continue
        fn = getattr(lib, name)
        args = fn.argtypes or []
        if not args or args[0] is not ctypes.c_void_p:
            continue
        call = [None if (a is ctypes.c_void_p or hasattr(a, "contents") or a.__name__.startswith("LP_")) else 0 for a in args]
        rc = fn(*call)
        if fn.restype in (ctypes.c_float,):
            assert rc == 0.0, name
        elif fn.restype in (ctypes.c_uint32,):
            assert rc == 0, name
        else:
            assert rc < 0, name
            assert len(lib.vgi_last_error()) > 0, name
        checked += 1
    assert checked > 50
    lib.vgi_destroy(None)
    lib.vgi_batch_destroy(None)
