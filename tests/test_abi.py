"""The C ABI library loads and exports every symbol include/vgi.h declares,
and only those; the workload library (include/vgi_workloads.h) likewise
(no GPU calls)."""
import ctypes
import os
import re

from conftest import ROOT, load_package


def declared_symbols(header="vgi.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(vgi_[a-z0-9_]+)\s*\(", text)))


def defined_symbols(path):
    """Global text symbols of a built shared library (nm -D --defined-only)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return sorted({ln.split()[-1] for ln in out.splitlines() if len(ln.split()) == 3 and ln.split()[1] == "T"})


def test_library_exports_every_declared_symbol():
    vgi = load_package()
    lib = vgi.library()
    names = declared_symbols()
    assert len(names) > 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # The Python layer binds exactly the declared surface.
    assert sorted(vgi.exported_symbols()) == names
    # The product exports no scene generator: those live in the workload library.
    exported = [n for n in defined_symbols(vgi.LIB_PATH) if n.startswith("vgi_")]
    assert exported == names, sorted(set(exported) ^ set(names))


def test_workload_library_is_separate():
    vgi = load_package()
    names = declared_symbols("vgi_workloads.h")
    assert names == sorted(vgi.workload_symbols())
    lib = vgi.workloads()
    assert all(hasattr(lib, n) for n in names)
    assert [n for n in defined_symbols(vgi.WORKLOADS_PATH) if n.startswith("vgi_")] == names


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
        if name in ("vgi_last_error", "vgi_destroy", "vgi_batch_destroy"):
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
