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
